// unet.hip -- the learned K-step jump (include/ppf_unet.h): a 3-level 3-D
// convolutional U-Net with hand-written forward, backward and Adam kernels.
//
// Convolutions are implicit GEMMs on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16: A 16x32 bf16, B 32x16 bf16, C 16x16 f32).  For a
// 3x3x3 periodic convolution Y[v][co] = sum_tap sum_ci X[v+s_tap][ci] *
// W[tap][co][ci], the M dimension is sites (16 per wave), N is output
// channels and K runs over (tap, input channel).  Activations are stored
// channels-last in BF16, so each lane's A operand (8 consecutive channels of
// one site) and B operand (8 consecutive input channels of one weight row)
// are single 16-byte loads.  The input gradient is the same convolution
// with the weights flipped in space and transposed in channels; the weight
// gradient contracts over sites (K = sites) and writes per-split partial
// sums that a second pass adds in a fixed order.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "ppf_host.h"
#include "ppf_unet.h"
#include "k_misc.h"

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr float LEAK = 0.1f;

// The convolution kernels run the MFMA as C^T = W X^T: the weight fragment
// is the A operand (rows = output channels) and the input fragment the B
// operand (columns = sites), so lane (l % 16, l / 16) ends with channels
// 4(l/16) .. 4(l/16)+3 of site l % 16, four consecutive values of one
// channels-last row, written with one 8-byte (bf16) or 16-byte (fp32) store.
// MODE 0 outputs: bf16(leaky(bias + sum)), the accumulator starting at the
// bias (a load of the bias in the store loop would make each store wait for
// the stores before it: vmcnt counts both).
__device__ __forceinline__ f32x4 bias4(const float *bias, int c0)
{
    return f32x4{bias[c0], bias[c0 + 1], bias[c0 + 2], bias[c0 + 3]};
}

__device__ __forceinline__ bf16x4 leaky4(f32x4 v)
{
    bf16x4 o;
#pragma unroll
    for (int e = 0; e < 4; e++)
        o[e] = (bf16)(v[e] > 0.f ? v[e] : LEAK * v[e]);
    return o;
}

constexpr int NCONV = 10;

// An empty asm that reads v: the compiler must wait for the load of v here.
// Before a loop this keeps a wait for loop-invariant operands (weights, a
// bias) out of the loop body, where vmcnt, counting in issue order, would
// also wait for every load and store issued after them.
__device__ __forceinline__ void settle(f32x4 v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void settle(bf16x8 v) { settle(__builtin_bit_cast(f32x4, v)); }

__device__ __forceinline__ int wrapi(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// Any coordinate to [0, n): a staged brick can be larger than the grid (the
// coarse levels of small inputs), so its halo may lie several periods out.
__device__ __forceinline__ int wrapm(int i, int n)
{
    const int r = i % n;
    return r < 0 ? r + n : r;
}

// Site index -> (b, z, y, x) for a [B][D][H][W] grid.
struct Coord {
    int b, z, y, x;
};

__device__ __forceinline__ Coord coord_of(size_t v, size_t V, int H, int W)
{
    Coord c;
    c.b = (int)(v / V);
    const size_t r = v - (size_t)c.b * V;
    const size_t hw = (size_t)H * W;
    c.z = (int)(r / hw);
    const size_t r2 = r - (size_t)c.z * hw;
    c.y = (int)(r2 / W);
    c.x = (int)(r2 - (size_t)c.y * W);
    return c;
}

// ----------------------------------------------------------------------------
// Convolution: MODE 0 stores bf16(leaky(acc + bias)); MODE 1 stores acc as
// FP32; MODE 2 adds acc into FP32.  Each wave computes 16 sites x NJ*16
// output channels; a workgroup is 4 waves (64 consecutive sites).  Cin is a
// multiple of 32, CoutW (the padded weight rows) a multiple of 16.
// ----------------------------------------------------------------------------
template <int NJ, int MODE, int TAPS>
__global__ __launch_bounds__(256) void k_conv(const bf16 *__restrict__ X, int ldx, int xoff,
                                              int Cin, const bf16 *__restrict__ Wt, int CoutW,
                                              const float *__restrict__ bias, int Cout,
                                              bf16 *__restrict__ Yb, float *__restrict__ Yf,
                                              int ldy, int yoff, int B, int D, int H, int Wd)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t V = (size_t)D * H * Wd, total = V * B;
    const size_t vbase = ((size_t)blockIdx.x * 4 + wave) * 16;
    const int n0 = blockIdx.y * (NJ * 16);
    const size_t va = vbase + (lane & 15);
    const bool va_ok = va < total;
    const Coord c = coord_of(va_ok ? va : 0, V, H, Wd);
    const int kq = (lane >> 4) * 8;
    f32x4 acc[NJ];
    for (int j = 0; j < NJ; j++)
        acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < TAPS; ++tap) {
        int dz = 0, dy = 0, dx = 0;
        if (TAPS == 27) {
            dz = tap / 9 - 1;
            dy = (tap / 3) % 3 - 1;
            dx = tap % 3 - 1;
        }
        const size_t nv = (size_t)c.b * V +
                          ((size_t)wrapi(c.z + dz, D) * H + wrapi(c.y + dy, H)) * Wd +
                          wrapi(c.x + dx, Wd);
        const bf16 *xa = X + nv * ldx + xoff + kq;
        const bf16 *wb = Wt + ((size_t)tap * CoutW + n0 + (lane & 15)) * Cin + kq;
        for (int kc = 0; kc < Cin; kc += 32) {
            bf16x8 a;
            if (va_ok)
                a = *(const bf16x8 *)(xa + kc);
            else
                for (int j = 0; j < 8; j++)
                    a[j] = (bf16)0.0f;
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                if (n0 + 16 * j < CoutW) {
                    const bf16x8 bw = *(const bf16x8 *)(wb + (size_t)16 * j * Cin + kc);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw, acc[j], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int ch = n0 + 16 * j + (lane & 15);
        if (ch >= Cout)
            continue;
        for (int r = 0; r < 4; r++) {
            const size_t vox = vbase + (lane >> 4) * 4 + r;
            if (vox >= total)
                continue;
            float v = acc[j][r];
            const size_t o = vox * ldy + yoff + ch;
            if (MODE == 0) {
                v += bias[ch];
                v = v > 0.f ? v : LEAK * v;
                Yb[o] = (bf16)v;
            } else if (MODE == 1) {
                Yf[o] = v;
            } else {
                Yf[o] += v;
            }
        }
    }
}

// ----------------------------------------------------------------------------
// LDS-tiled convolution, same contract as k_conv.  A workgroup computes a
// brick of 16 (x) x 8 (y) sites of one z plane for 64 output channels.  Per
// chunk of 32 input channels it stages the brick's halo (3 x 10 x 18 sites)
// in LDS, then per row of taps (dz, dy) the weights of the 3 taps dx; each
// wave multiplies 2 site rows (2 M tiles) by 4 N tiles: 8 MFMAs per 6 LDS
// fragment reads.  Site and weight rows are padded to 40 bf16 (80 bytes) so
// the 16 lanes of a fragment read hit distinct banks.
// ----------------------------------------------------------------------------
constexpr int CL_BX = 16, CL_BY = 8, CL_HX = CL_BX + 2, CL_HY = CL_BY + 2;
constexpr int CL_HALO = 3 * CL_HY * CL_HX;  // 540 sites
constexpr int CL_PAD = 40;                   // bf16 per staged site / weight row
constexpr int CL_CT = 64;                    // output channels per workgroup

template <int MODE>
__global__ __launch_bounds__(256) void k_conv_lds(const bf16 *__restrict__ X, int ldx, int xoff,
                                                  int Cin, const bf16 *__restrict__ Wt, int CoutW,
                                                  const float *__restrict__ bias, int Cout,
                                                  bf16 *__restrict__ Yb, float *__restrict__ Yf,
                                                  int ldy, int yoff, int B, int D, int H, int Wd)
{
    __shared__ __attribute__((aligned(16))) bf16 sin[CL_HALO * CL_PAD];
    __shared__ __attribute__((aligned(16))) bf16 sw[3 * CL_CT * CL_PAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbx = (Wd + CL_BX - 1) / CL_BX, nby = (H + CL_BY - 1) / CL_BY;
    int r = blockIdx.x;
    const int bx = r % nbx;
    r /= nbx;
    const int by = r % nby;
    r /= nby;
    const int z = r % D;
    const int b = r / D;
    const int x0 = bx * CL_BX, y0 = by * CL_BY;
    const int n0 = blockIdx.y * CL_CT;
    const size_t V = (size_t)D * H * Wd;
    const int ntile = min(4, (CoutW - n0 + 15) / 16);  // N tiles that exist
    const int kq = (lane >> 4) * 8;
    f32x4 acc[2][4];
    for (int m = 0; m < 2; m++)
        for (int j = 0; j < 4; j++)
            acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Cin; kc += 32) {
        __syncthreads();  // the previous chunk's reads of sin are done
        for (int i = tid; i < CL_HALO * 4; i += 256) {
            const int s = i >> 2, part = i & 3;
            const int hz = s / (CL_HY * CL_HX), hy = (s / CL_HX) % CL_HY, hx = s % CL_HX;
            const size_t nv = (size_t)b * V +
                              ((size_t)wrapm(z + hz - 1, D) * H + wrapm(y0 + hy - 1, H)) * Wd +
                              wrapm(x0 + hx - 1, Wd);
            *(bf16x8 *)(sin + s * CL_PAD + part * 8) =
                *(const bf16x8 *)(X + nv * ldx + xoff + kc + part * 8);
        }
        for (int g = 0; g < 9; g++) {
            const int dz = g / 3 - 1, dy = g % 3 - 1;
            __syncthreads();  // sin staged / the previous group's reads of sw done
            for (int i = tid; i < 3 * CL_CT * 4; i += 256) {
                const int part = i & 3, row = i >> 2;  // row = dxi*64 + co
                const int dxi = row / CL_CT, co = n0 + row % CL_CT;
                bf16x8 v;
                if (co < CoutW)
                    v = *(const bf16x8 *)(Wt + ((size_t)(g * 3 + dxi) * CoutW + co) * Cin + kc +
                                          part * 8);
                else
                    for (int e = 0; e < 8; e++)
                        v[e] = (bf16)0.0f;
                *(bf16x8 *)(sw + row * CL_PAD + part * 8) = v;
            }
            __syncthreads();
            for (int dxi = 0; dxi < 3; dxi++) {
                bf16x8 a[2];
                for (int m = 0; m < 2; m++) {
                    const int hy = 2 * wave + m + dy + 1, hx = (lane & 15) + dxi;
                    a[m] = *(const bf16x8 *)(sin + (((dz + 1) * CL_HY + hy) * CL_HX + hx) * CL_PAD +
                                             kq);
                }
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (j < ntile) {
                        const bf16x8 w =
                            *(const bf16x8 *)(sw + (dxi * CL_CT + j * 16 + (lane & 15)) * CL_PAD + kq);
                        acc[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], w, acc[0][j], 0, 0, 0);
                        acc[1][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], w, acc[1][j], 0, 0, 0);
                    }
                }
            }
        }
    }
    // lane holds C[site 4*(lane>>4)+rr][channel lane&15] of each (m, j) tile
    for (int m = 0; m < 2; m++) {
        const int y = y0 + 2 * wave + m;
        if (y >= H)
            continue;
        for (int j = 0; j < 4; j++) {
            const int ch = n0 + j * 16 + (lane & 15);
            if (j >= ntile || ch >= Cout)
                continue;
            for (int rr = 0; rr < 4; rr++) {
                const int x = x0 + (lane >> 4) * 4 + rr;
                if (x >= Wd)
                    continue;
                const size_t vox = (size_t)b * V + ((size_t)z * H + y) * Wd + x;
                float v = acc[m][j][rr];
                const size_t o = vox * ldy + yoff + ch;
                if (MODE == 0) {
                    v += bias[ch];
                    v = v > 0.f ? v : LEAK * v;
                    Yb[o] = (bf16)v;
                } else if (MODE == 1) {
                    Yf[o] = v;
                } else {
                    Yf[o] += v;
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Row-blocked convolution, same contract as k_conv; no LDS, no barriers.
// A wave owns 16 consecutive x sites of R consecutive rows (one z plane) and
// NJ*16 output channels.  For one (dz, dx) and 32 input channels it loads
// the R+2 input rows y0-1 .. y0+R as A fragments; row r+dy serves output row
// r for tap dy, so each row is loaded once for all three dy taps: per (dz,
// dx) group R+2 A and 3*NJ B fragments feed 3*R*NJ MFMAs, and the next
// group's fragments load while they run.  Every
// coordinate is wrapped to the grid (a tile may be larger than a coarse
// grid); sites outside the grid are computed on wrapped data and not stored.
// Waves are ordered N block, x tile, y tile, z, batch, so the four waves of
// a workgroup share their input rows in L1.
// ----------------------------------------------------------------------------
// UP: the first `cup` input channels are the nearest x2 upsampling of Xc
// ([B][D/2][H/2][W/2][ldc]), read at the coarse site (z/2, y/2, x/2) of each
// (wrapped) fine neighbour instead of from a materialized copy in X; the
// values, and so the sums, are the same as with the copy.  Input channel
// c >= cup is then X's element xoff + c - cup of a site (the skip channels
// alone, or after room left for the upsampled ones).
template <int R, int NJ, int MODE, int WPE, bool UP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_conv_rows(
    const bf16 *__restrict__ X, int ldx, int xoff, int Cin, const bf16 *__restrict__ Wt, int CoutW,
    const float *__restrict__ bias, int Cout, bf16 *__restrict__ Yb, float *__restrict__ Yf,
    int ldy, int yoff, int B, int D, int H, int Wd, int nnb, int nxt, int nyt,
    const bf16 *__restrict__ Xc, int ldc, int cup)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform
    long g = (long)blockIdx.x * 4 + wave;
    const int nb = (int)(g % nnb);
    g /= nnb;
    const int xt = (int)(g % nxt);
    g /= nxt;
    const int yt = (int)(g % nyt);
    g /= nyt;
    const int z = (int)(g % D);
    const int b = (int)(g / D);
    if (b >= B)
        return;  // whole wave; the kernel has no barrier
    const int xl = lane & 15, q = lane >> 4;
    const int n0 = nb * NJ * 16;
    const int y0 = yt * R;
    const int rowlen = Wd * ldx;  // elements per grid row; a plane is < 2^31 elements
    // per-lane offsets (elements) for dx = -1, 0, +1, with the channel quad;
    // everything else in an address is uniform (scalar base + vector offset)
    int xo[3];  // bytes
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
        xo[dx] = (wrapm(xt * 16 + xl + dx - 1, Wd) * ldx + xoff + 8 * q) * 2;
    // Loads are buffer loads: a descriptor per source in scalar registers, a
    // per-lane byte offset (VGPR) and a uniform byte offset (SGPR), so no
    // 64-bit address is kept per load.  Each descriptor's extent is its
    // source's true size (a plane, the weight rows); a load outside it
    // would read zeros, not fault.
    // planes z-1, z, z+1 and rows y0-1 .. y0+R
    const int planebytes = H * rowlen * 2;
    __amdgpu_buffer_rsrc_t pl[3];
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
        pl[dz] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(X + ((size_t)b * D + wrapm(z + dz - 1, D)) * H * (size_t)rowlen), (short)0,
            planebytes, 0x00020000);
    int ro[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++)
        ro[r] = wrapm(y0 + r - 1, H) * rowlen;
    // the coarse source of the upsampled channels (UP)
    int xoc[3] = {0, 0, 0}, roc[R + 2];
    __amdgpu_buffer_rsrc_t plc[3];
    if (UP) {
        const int Hc = H / 2, rowlen_c = (Wd / 2) * ldc;
#pragma unroll
        for (int dx = 0; dx < 3; dx++)
            xoc[dx] = ((wrapm(xt * 16 + xl + dx - 1, Wd) >> 1) * ldc + 8 * q) * 2;
#pragma unroll
        for (int dz = 0; dz < 3; dz++)
            plc[dz] = __builtin_amdgcn_make_buffer_rsrc(
                (void *)(Xc + ((size_t)b * (D / 2) + (wrapm(z + dz - 1, D) >> 1)) * Hc *
                                  (size_t)rowlen_c),
                (short)0, Hc * rowlen_c * 2, 0x00020000);
#pragma unroll
        for (int r = 0; r < R + 2; r++)
            roc[r] = (wrapm(y0 + r - 1, H) >> 1) * rowlen_c;
    }
    // lane's weight row in N tile j: Wt[tap][n0+16j+xl][kc+8q]; MODE 0 (the
    // forward) reads the fragment-contiguous copy Wt[tap][kc/32][n0+16j+xl][8q]
    // (k_prep_w).  CoutW is a multiple of NJ*16, so every N tile exists.
    constexpr bool WCH = MODE == 0;
    const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Wt + (size_t)n0 * (WCH ? 32 : Cin)), (short)0, (27 * CoutW * Cin - n0 * (WCH ? 32 : Cin)) * 2,
        0x00020000);
    const int wo = (xl * (WCH ? 32 : Cin) + 8 * q) * 2;
    const int tapstride = CoutW * Cin, jstride = 16 * (WCH ? 32 : Cin);
    // element offset of chunk kc's fragments of a tap (N tile 0)
    auto wk = [&](int tap, int kc) { return WCH ? tap * tapstride + kc * CoutW : tap * tapstride + kc; };
    // MODE 0: the accumulators start at the bias of the lane's four output
    // channels
    f32x4 acc[R][NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int c0 = n0 + 16 * j + 4 * q;
        const f32x4 bj = (MODE == 0 && c0 < Cout) ? bias4(bias, c0) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < R; r++)
            acc[r][j] = bj;
    }
    // Fragments of one (32-channel chunk, dz, dx) group.
    auto ld = [&](bf16x8(&a)[R + 2], bf16x8(&w)[3][NJ], int kc, int grp) {
        const int dz = grp / 3, dx = grp % 3;
        if (UP && kc < cup) {  // uniform; the same number of loads either way
#pragma unroll
            for (int r = 0; r < R + 2; r++)
                a[r] = __builtin_bit_cast(
                    bf16x8,
                    __builtin_amdgcn_raw_buffer_load_b128(plc[dz], xoc[dx], (roc[r] + kc) * 2, 0));
        } else {
#pragma unroll
            for (int r = 0; r < R + 2; r++)
                a[r] = __builtin_bit_cast(
                    bf16x8,
                    __builtin_amdgcn_raw_buffer_load_b128(pl[dz], xo[dx], (ro[r] + kc - (UP ? cup : 0)) * 2, 0));
        }
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int j = 0; j < NJ; j++)
                w[dy][j] = __builtin_bit_cast(
                    bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                wr, wo, (wk(dz * 9 + dy * 3 + dx, kc) + j * jstride) * 2, 0));
    };
    auto mm = [&](const bf16x8(&a)[R + 2], const bf16x8(&w)[3][NJ]) {
#pragma unroll
        for (int dy = 0; dy < 3; dy++)
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[dy][j], a[r + dy],
                                                                        acc[r][j], 0, 0, 0);
    };
    // Two register buffers.  Each group's MFMAs run while the next group's
    // loads are in flight; scheduling fences keep the loads ahead of the
    // MFMAs (left alone, the compiler moves each load next to its use).
    // A chunk has 9 groups, so the buffers swap roles from one chunk to the
    // next and the loop body covers two chunks.  Past the last chunk the
    // prefetch re-reads the last chunk (in range, unused).
    bf16x8 a0[R + 2], w0[3][NJ], a1[R + 2], w1[3][NJ];
    const int klast = Cin - 32;
    ld(a0, w0, 0, 0);
    for (int kc = 0; kc < Cin; kc += 64) {
        const int k1 = min(kc + 32, klast), k2 = min(kc + 64, klast);
#pragma unroll
        for (int grp = 0; grp < 9; grp++) {
            if (grp % 2 == 0) {
                ld(a1, w1, grp < 8 ? kc : k1, grp < 8 ? grp + 1 : 0);
                __builtin_amdgcn_sched_barrier(0);
                mm(a0, w0);
            } else {
                ld(a0, w0, kc, grp + 1);
                __builtin_amdgcn_sched_barrier(0);
                mm(a1, w1);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (kc + 32 >= Cin)
            break;
#pragma unroll
        for (int grp = 0; grp < 9; grp++) {
            if (grp % 2 == 0) {
                ld(a0, w0, grp < 8 ? k1 : k2, grp < 8 ? grp + 1 : 0);
                __builtin_amdgcn_sched_barrier(0);
                mm(a1, w1);
            } else {
                ld(a1, w1, k1, grp + 1);
                __builtin_amdgcn_sched_barrier(0);
                mm(a0, w0);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // lane holds channels n0+16j+4q .. +3 of site (xt*16+xl, y0+r) in acc[r][j]
    const int x = xt * 16 + xl;
    if (x >= Wd)
        return;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + r;
        if (y >= H)
            break;
        const size_t vox = (((size_t)b * D + z) * H + y) * Wd + x;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int c0 = n0 + 16 * j + 4 * q;
            if (c0 >= Cout)
                continue;
            const size_t o = vox * ldy + yoff + c0;
            if (MODE == 0) {
                *(bf16x4 *)(Yb + o) = leaky4(acc[r][j]);
            } else if (MODE == 1) {
                *(f32x4 *)(Yf + o) = acc[r][j];
            } else {
                *(f32x4 *)(Yf + o) += acc[r][j];
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Decoder convolution over [nearest x2 upsampling of Xc | skip channels of X],
// polyphase form (inference).  A fine site x = 2X + p (per dimension) sees
// its neighbours x-1, x, x+1 of an upsampled tensor at the coarse offsets
// floor((p + d) / 2): {-1, 0, 0} for p = 0, {0, 0, +1} for p = 1.  So for
// each output parity (pz, py, px) the 27-tap convolution of the upsampled
// channels is a 2x2x2 convolution of the coarse tensor, whose weights Wp
// sum the fine taps that land on the same coarse offset (k_prep_wpoly, from
// the FP32 weights, rounded to BF16 once): 8 taps instead of 27.  The skip
// channels (input channels cup .. Cin-1: X's elements xoff .. xoff + Cin -
// cup - 1 of a site) take the ordinary 27 taps.
// A wave owns the 16 sites x = 32 xt + 2 l + px (l = lane % 16) of one
// parity px, R rows (R even, y0 even: row r has py = r & 1) of one z plane,
// and NJ*16 output channels.  Per 32 upsampled channels and (oz, ox) coarse
// offset, the R/2 + 2 coarse rows serve all R rows: row r reads coarse rows
// (r >> 1) + py + oy (oy = 0, 1) with the weights of (py, oy).  Output as
// k_conv_rows MODE 0.  D, H, W even.
// ----------------------------------------------------------------------------
// UP_ONLY: the upsampled channels only (bias + their sum), stored as FP32
// partial sums in Yf [site][ldy] for k_conv_zs64's partial-input mode.
template <int R, int NJ, int WPE, bool UP_ONLY = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_conv_poly(
    const bf16 *__restrict__ X, int ldx, int xoff, int Cin, const bf16 *__restrict__ Wt, int CoutW,
    const bf16 *__restrict__ Wp, const float *__restrict__ bias, int Cout, bf16 *__restrict__ Yb,
    int ldy, int B, int D, int H, int Wd, int nnb, int nxt2, int nyt, const bf16 *__restrict__ Xc,
    int ldc, int cup, float *__restrict__ Yf = nullptr)
{
    static_assert(R % 2 == 0, "rows come in parity pairs");
    constexpr int RC = R / 2 + 2;  // coarse rows per wave
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    long g = (long)blockIdx.x * 4 + wave;
    const int nb = (int)(g % nnb);
    g /= nnb;
    const int xt2 = (int)(g % nxt2);
    g /= nxt2;
    const int yt = (int)(g % nyt);
    g /= nyt;
    const int z = (int)(g % D);
    const int b = (int)(g / D);
    if (b >= B)
        return;  // whole wave; the kernel has no barrier
    const int px = xt2 & 1, xt = xt2 >> 1, pz = z & 1;
    const int xl = lane & 15, q = lane >> 4;
    const int n0 = nb * NJ * 16;
    const int y0 = yt * R;
    const int Dc = D / 2, Hc = H / 2, Wc = Wd / 2;
    // skip channels: fine planes z-1 .. z+1, rows y0-1 .. y0+R, the lane's
    // site x = 32 xt + 2 xl + px and its x neighbours
    const int rowlen = Wd * ldx;
    int xo[3];
#pragma unroll
    for (int dx = 0; dx < 3; dx++)
        xo[dx] = (wrapm(32 * xt + 2 * xl + px + dx - 1, Wd) * ldx + 8 * q) * 2;
    const int planebytes = H * rowlen * 2;
    __amdgpu_buffer_rsrc_t pl[3];
#pragma unroll
    for (int dz = 0; dz < 3; dz++)
        pl[dz] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(X + ((size_t)b * D + wrapm(z + dz - 1, D)) * H * (size_t)rowlen), (short)0,
            planebytes, 0x00020000);
    int ro[R + 2];
#pragma unroll
    for (int r = 0; r < R + 2; r++)
        ro[r] = wrapm(y0 + r - 1, H) * rowlen;
    // upsampled channels: coarse planes (z >> 1) + pz - 1 + oz, rows
    // y0/2 - 1 + cr, x (16 xt + xl) + px - 1 + ox
    const int rowlen_c = Wc * ldc;
    int xoc[2];
#pragma unroll
    for (int ox = 0; ox < 2; ox++)
        xoc[ox] = (wrapm(16 * xt + xl + px - 1 + ox, Wc) * ldc + 8 * q) * 2;
    __amdgpu_buffer_rsrc_t plc[2];
#pragma unroll
    for (int oz = 0; oz < 2; oz++)
        plc[oz] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(Xc + ((size_t)b * Dc + wrapm((z >> 1) + pz - 1 + oz, Dc)) * Hc *
                              (size_t)rowlen_c),
            (short)0, Hc * rowlen_c * 2, 0x00020000);
    int roc[RC];
#pragma unroll
    for (int cr = 0; cr < RC; cr++)
        roc[cr] = wrapm(y0 / 2 - 1 + cr, Hc) * rowlen_c;
    // weights, fragment-contiguous: Wt[tap][kc/32][n0+16j+xl][8q] (skip,
    // k_prep_w), Wp[par][grp][kc/32][n0+16j+xl][8q] (upsampled, par = 4 pz +
    // 2 py + px, grp = 4 oz + 2 oy + ox)
    const __amdgpu_buffer_rsrc_t wr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(Wt + (size_t)n0 * 32), (short)0, (27 * CoutW * Cin - n0 * 32) * 2, 0x00020000);
    const int wo = (xl * 32 + 8 * q) * 2;
    const int tapstride = CoutW * Cin, jstride = 16 * 32;
    const __amdgpu_buffer_rsrc_t wpr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)Wp, (short)0, 64 * CoutW * cup * 2, 0x00020000);
    const int wpo = (xl * 32 + 8 * q) * 2;
    // accumulators start at the bias of the lane's four output channels
    f32x4 acc[R][NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int c0 = n0 + 16 * j + 4 * q;
        const f32x4 bj = c0 < Cout ? bias4(bias, c0) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < R; r++)
            acc[r][j] = bj;
    }
    // upsampled channels: per 32-channel chunk, four (oz, ox) groups, two
    // register buffers (the next group's loads in flight during the MFMAs)
    {
        bf16x8 c0[RC], q0[2][2][NJ], c1[RC], q1[2][2][NJ];
        auto ldp = [&](bf16x8(&c)[RC], bf16x8(&w)[2][2][NJ], int kc, int grp) {
            const int oz = grp >> 1, ox = grp & 1;
#pragma unroll
            for (int cr = 0; cr < RC; cr++)
                c[cr] = __builtin_bit_cast(
                    bf16x8,
                    __builtin_amdgcn_raw_buffer_load_b128(plc[oz], xoc[ox], (roc[cr] + kc) * 2, 0));
#pragma unroll
            for (int py = 0; py < 2; py++)
#pragma unroll
                for (int oy = 0; oy < 2; oy++)
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        w[py][oy][j] = __builtin_bit_cast(
                            bf16x8,
                            __builtin_amdgcn_raw_buffer_load_b128(
                                wpr, wpo,
                                (((((4 * pz + 2 * py + px) * 8 + 4 * oz + 2 * oy + ox) * (cup / 32) +
                                   kc / 32) * CoutW + n0 + 16 * j) * 32) * 2,
                                0));
        };
        // oy outermost: each accumulator still takes oy = 0 then 1, and R*NJ
        // independent MFMAs separate the two
        auto mmp = [&](const bf16x8(&c)[RC], const bf16x8(&w)[2][2][NJ]) {
#pragma unroll
            for (int oy = 0; oy < 2; oy++)
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            w[r & 1][oy][j], c[(r >> 1) + (r & 1) + oy], acc[r][j], 0, 0, 0);
        };
        // past the last chunk the prefetch re-reads the last chunk (unused)
        const int klast = cup - 32;
        ldp(c0, q0, 0, 0);
        for (int kc = 0; kc < cup; kc += 32) {
            const int kn = min(kc + 32, klast);
            ldp(c1, q1, kc, 1);
            __builtin_amdgcn_sched_barrier(0);
            mmp(c0, q0);
            __builtin_amdgcn_sched_barrier(0);
            ldp(c0, q0, kc, 2);
            __builtin_amdgcn_sched_barrier(0);
            mmp(c1, q1);
            __builtin_amdgcn_sched_barrier(0);
            ldp(c1, q1, kc, 3);
            __builtin_amdgcn_sched_barrier(0);
            mmp(c0, q0);
            __builtin_amdgcn_sched_barrier(0);
            ldp(c0, q0, kn, 0);
            __builtin_amdgcn_sched_barrier(0);
            mmp(c1, q1);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // skip channels: the 27 taps, as k_conv_rows (9 (dz, dx) groups per
    // chunk, R + 2 rows serve the three dy taps)
    if (!UP_ONLY && cup < Cin) {
        bf16x8 a0[R + 2], w0[3][NJ], a1[R + 2], w1[3][NJ];
        auto ld = [&](bf16x8(&a)[R + 2], bf16x8(&w)[3][NJ], int kc, int grp) {
            const int dz = grp / 3, dx = grp % 3;
#pragma unroll
            for (int r = 0; r < R + 2; r++)
                a[r] = __builtin_bit_cast(
                    bf16x8,
                    __builtin_amdgcn_raw_buffer_load_b128(pl[dz], xo[dx], (ro[r] + xoff + kc - cup) * 2, 0));
#pragma unroll
            for (int dy = 0; dy < 3; dy++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    w[dy][j] = __builtin_bit_cast(
                        bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                    wr, wo,
                                    ((dz * 9 + dy * 3 + dx) * tapstride + kc * CoutW + j * jstride) * 2,
                                    0));
        };
        auto mm = [&](const bf16x8(&a)[R + 2], const bf16x8(&w)[3][NJ]) {
#pragma unroll
            for (int dy = 0; dy < 3; dy++)
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int j = 0; j < NJ; j++)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[dy][j], a[r + dy],
                                                                            acc[r][j], 0, 0, 0);
        };
        const int klast = Cin - 32;
        ld(a0, w0, cup, 0);
        for (int kc = cup; kc < Cin; kc += 64) {
            const int k1 = min(kc + 32, klast), k2 = min(kc + 64, klast);
#pragma unroll
            for (int grp = 0; grp < 9; grp++) {
                if (grp % 2 == 0) {
                    ld(a1, w1, grp < 8 ? kc : k1, grp < 8 ? grp + 1 : 0);
                    __builtin_amdgcn_sched_barrier(0);
                    mm(a0, w0);
                } else {
                    ld(a0, w0, kc, grp + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    mm(a1, w1);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            if (kc + 32 >= Cin)
                break;
#pragma unroll
            for (int grp = 0; grp < 9; grp++) {
                if (grp % 2 == 0) {
                    ld(a0, w0, grp < 8 ? k1 : k2, grp < 8 ? grp + 1 : 0);
                    __builtin_amdgcn_sched_barrier(0);
                    mm(a1, w1);
                } else {
                    ld(a1, w1, k1, grp + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    mm(a0, w0);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    }
    // lane holds channels n0+16j+4q .. +3 of site (32 xt + 2 xl + px, y0+r)
    const int x = 32 * xt + 2 * xl + px;
    if (x >= Wd)
        return;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int y = y0 + r;
        if (y >= H)
            break;
        const size_t vox = (((size_t)b * D + z) * H + y) * Wd + x;
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int c0 = n0 + 16 * j + 4 * q;
            if (c0 < Cout) {
                if (UP_ONLY)
                    *(f32x4 *)(Yf + vox * ldy + c0) = acc[r][j];
                else
                    *(bf16x4 *)(Yb + vox * ldy + c0) = leaky4(acc[r][j]);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// k_conv_poly's upsampled-channel part alone (its UP_ONLY mode: bias + the
// polyphase sum into FP32 partial sums Yf [site][ldy]) for 64 output
// channels, with the weights read once per workgroup.  The four waves of a
// workgroup own four consecutive 16-row tiles of the same fine plane z and
// x parity px, so they multiply with the same weights: per step (32-channel
// chunk kc, coarse offsets oz, ox) the 16 fragments of (py, oy, j), 16 KB,
// are loaded from global memory by the whole workgroup into one of two LDS
// buffers while the previous step's MFMAs run, and every wave reads them
// from there (k_conv_poly has each wave stream them from the cache, which
// bounds it by the cache bandwidth).  The coarse rows each wave multiplies
// are loaded as in k_conv_poly, one step ahead.  Same MFMAs in the same
// order per accumulator as k_conv_poly<16, 4, WPE, true>, so the partial
// sums are equal bit for bit.  NW waves of R rows (NW R = 64: 4 x 16 rows
// at one wave per SIMD, 8 x 8 at two).  W % 32 == 0, H % 64 == 0, D even.
// ----------------------------------------------------------------------------
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4))) void k_conv_polyup(
    const bf16 *__restrict__ Wp, const float *__restrict__ bias, int ldy, int B, int D, int H,
    int Wd, const bf16 *__restrict__ Xc, int ldc, int cup, float *__restrict__ Yf)
{
    static_assert(R * NW == 64 && R % 2 == 0, "64 rows per workgroup");
    constexpr int NJ = 4, RC = R / 2 + 2, CoutW = 64;
    // per thread: WB 16-byte pieces of one 4 KB weight block
    constexpr int TPB = 16 * NW, WB = 16 / NW;
    __shared__ __attribute__((aligned(16))) bf16 wsh[2][4][CoutW * 32];  // 32 KB
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nxt2 = Wd / 16, nytg = H / (NW * R);
    unsigned g = blockIdx.x;
    const int ytg = (int)(g % nytg);
    g /= nytg;
    const int xt2 = (int)(g % nxt2);
    g /= nxt2;
    const int z = (int)(g % D);
    const int b = (int)(g / D);
    const int px = xt2 & 1, xt = xt2 >> 1, pz = z & 1;
    const int xl = lane & 15, q = lane >> 4;
    const int y0 = (NW * ytg + wave) * R;
    const int Dc = D / 2, Hc = H / 2, Wc = Wd / 2;
    // coarse planes (z >> 1) + pz - 1 + oz, rows y0/2 - 1 + cr, x (16 xt +
    // xl) + px - 1 + ox
    const int rowlen_c = Wc * ldc;
    int xoc[2];
#pragma unroll
    for (int ox = 0; ox < 2; ox++)
        xoc[ox] = (wrapm(16 * xt + xl + px - 1 + ox, Wc) * ldc + 8 * q) * 2;
    __amdgpu_buffer_rsrc_t plc[2];
#pragma unroll
    for (int oz = 0; oz < 2; oz++)
        plc[oz] = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(Xc + ((size_t)b * Dc + wrapm((z >> 1) + pz - 1 + oz, Dc)) * Hc *
                              (size_t)rowlen_c),
            (short)0, Hc * rowlen_c * 2, 0x00020000);
    int roc[RC];
#pragma unroll
    for (int cr = 0; cr < RC; cr++)
        roc[cr] = wrapm(y0 / 2 - 1 + cr, Hc) * rowlen_c;
    // weights Wp[par][grp][kc/32][64][32] (par = 4 pz + 2 py + px, grp = 4 oz
    // + 2 oy + ox): step s = (kc, oz, ox) needs the four 4 KB blocks (py, oy);
    // thread t copies 16 WB bytes of block t / TPB
    const __amdgpu_buffer_rsrc_t wpr = __builtin_amdgcn_make_buffer_rsrc(
        (void *)Wp, (short)0, 64 * CoutW * cup * 2, 0x00020000);
    const int tb = threadIdx.x / TPB, tpy = tb >> 1, toy = tb & 1, tof = (threadIdx.x % TPB) * 8 * WB;
    const int nsteps = cup / 32 * 4;
    auto wsrc = [&](int s) {
        const int kc = s >> 2, oz = (s >> 1) & 1, ox = s & 1;
        return ((((4 * pz + 2 * tpy + px) * 8 + 4 * oz + 2 * toy + ox) * (cup / 32) + kc) * CoutW * 32 +
                tof) * 2;
    };
    bf16x8 wr[WB];
    auto wload = [&](int s) {
        const int off = wsrc(s);
#pragma unroll
        for (int i = 0; i < WB; i++)
            wr[i] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(wpr, off + 16 * i, 0, 0));
    };
    auto wstore = [&](int buf) {
#pragma unroll
        for (int i = 0; i < WB; i++)
            *(bf16x8 *)(&wsh[buf][tb][tof + 8 * i]) = wr[i];
    };
    auto cload = [&](bf16x8(&c)[RC], int s) {
        const int kc = (s >> 2) * 32, oz = (s >> 1) & 1, ox = s & 1;
#pragma unroll
        for (int cr = 0; cr < RC; cr++)
            c[cr] = __builtin_bit_cast(
                bf16x8, __builtin_amdgcn_raw_buffer_load_b128(plc[oz], xoc[ox], (roc[cr] + kc) * 2, 0));
    };
    f32x4 acc[R][NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const f32x4 bj = bias4(bias, 16 * j + 4 * q);
#pragma unroll
        for (int r = 0; r < R; r++)
            acc[r][j] = bj;
    }
    // one step: fragments from LDS buffer buf, then the MFMAs in
    // k_conv_poly's order (oy outermost)
    auto mm = [&](const bf16x8(&c)[RC], int buf) {
#pragma unroll
        for (int oy = 0; oy < 2; oy++) {
            bf16x8 w[2][NJ];
#pragma unroll
            for (int py = 0; py < 2; py++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    w[py][j] = *(const bf16x8 *)(&wsh[buf][2 * py + oy][(16 * j + xl) * 32 + 8 * q]);
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    acc[r][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        w[r & 1][j], c[(r >> 1) + (r & 1) + oy], acc[r][j], 0, 0, 0);
        }
    };
    bf16x8 c0[RC], c1[RC];
    wload(0);
    cload(c0, 0);
    wstore(0);
    __syncthreads();
    // steps in pairs (nsteps is a multiple of 4): even steps multiply c0
    // from buffer 0, odd ones c1 from buffer 1; each step first issues the
    // next step's loads
    // (past the last step the loads re-read the last step, unused)
    for (int s = 0; s < nsteps; s += 2) {
        wload(s + 1);
        cload(c1, s + 1);
        __builtin_amdgcn_sched_barrier(0);
        mm(c0, 0);
        __builtin_amdgcn_sched_barrier(0);
        wstore(1);
        __syncthreads();
        const int sn = min(s + 2, nsteps - 1);
        wload(sn);
        cload(c0, sn);
        __builtin_amdgcn_sched_barrier(0);
        mm(c1, 1);
        __builtin_amdgcn_sched_barrier(0);
        wstore(0);
        __syncthreads();
    }
    // lane holds channels 16j+4q .. +3 of site (32 xt + 2 xl + px, y0+r)
    const int x = 32 * xt + 2 * xl + px;
#pragma unroll
    for (int r = 0; r < R; r++) {
        const size_t vox = (((size_t)b * D + z) * H + y0 + r) * Wd + x;
#pragma unroll
        for (int j = 0; j < NJ; j++)
            *(f32x4 *)(Yf + vox * ldy + 16 * j + 4 * q) = acc[r][j];
    }
}

// LDS layout of the staged planes.  A fragment read (ds_read_b128) has lane
// (l % 16, l / 16) read 16 bytes at site l % 16 of a row, channel quad
// l / 16.  Measured on gfx950 (bank-conflict counter over site strides of
// 64 .. 272 bytes): it is conflict-free exactly when the site stride, in
// 16-byte units, is 2 mod 4 (96 and 160 bytes: none; 80, 112, 144, 176, 208,
// 240, 272: 4 extra cycles per read; 128: 12), i.e. the lanes of sites
// 0..7 and two channel quads are served together.  So staged sites are
// padded to 48 bf16 (32 channels) or 80 bf16 (64 channels) where the LDS
// allows it; k_conv_z32 keeps 40 (its 34 x 18 planes would not fit three
// times at 48, and a 28-row tile was slower than the conflicts, with 4 or 8
// waves).
constexpr int LDS_PAD32 = 48, LDS_PAD64 = 80;

// ----------------------------------------------------------------------------
// The 1x1x1 head (CB = 32 NK channels -> 3) of 16 sites on the matrix cores:
// C^T = Hw F^T, the FP32 head weights split into BF16 hi + lo parts (hi =
// bf16(w), lo = bf16(w - hi): w to 2^-17 relative, the products exact in the
// FP32 accumulator).  Whd (k_prep_head): [NK][2][16][32] BF16, rows 3..15
// zero.  f[k]: lane (l % 16, q = l / 16) holds channels 32 k + 8 q .. +7 of
// site l % 16.  The accumulator starts at the bias; lane (site, q = 0) ends
// with the site's three outputs in [0..2].  Every head (k_head_fwd,
// k_head_ctx and the fused epilogue of k_conv_zs32) runs this routine, so
// their outputs are bitwise equal.
// ----------------------------------------------------------------------------
template <int NK>
__device__ __forceinline__ f32x4 head_mfma(const bf16 *__restrict__ Whd, const float *__restrict__ hb,
                                           const bf16x8 (&f)[NK], int lane)
{
    const int xl = lane & 15, q = lane >> 4;
    const float b0 = hb[0], b1 = hb[1], b2 = hb[2];
    f32x4 c = q == 0 ? f32x4{b0, b1, b2, 0.f} : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < NK; k++)
#pragma unroll
        for (int h = 0; h < 2; h++)
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *(const bf16x8 *)(Whd + ((size_t)(2 * k + h) * 16 + xl) * 32 + 8 * q), f[k], c, 0, 0, 0);
    return c;
}

// Whd from the FP32 head weights hw [3][CB].
__global__ __launch_bounds__(256) void k_prep_head(const float *__restrict__ hw, int CB,
                                                   bf16 *__restrict__ Whd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;  // (k, h, row, c)
    if (i >= CB / 32 * 2 * 16 * 32)
        return;
    const int c = i & 31, row = (i >> 5) & 15, h = (i >> 9) & 1, k = i >> 10;
    const float w = row < 3 ? hw[row * CB + 32 * k + c] : 0.f;
    const bf16 hi = (bf16)w;
    Whd[i] = h == 0 ? hi : (bf16)(w - (float)hi);
}

// ----------------------------------------------------------------------------
// z-marching convolution for Cin = Cout = 32 (the level-1 layers), forward
// (MODE 0).  A workgroup owns a 16 (x) x 32 (y) tile and marches a chunk of
// z planes; wave w computes rows 8w .. 8w+7.  Each input plane is staged
// once in LDS (tile + one-site halo, 34 x 18 sites, 32 channels padded to
// 40) and serves the three output planes that need it; three slots hold
// planes z-1, z, z+1 while the next plane is loaded into registers.  All 27
// taps' weights (27 x 2 fragments) stay in registers.  Per output plane a
// wave runs 9 (dz, dx) groups of 10 LDS fragment reads and 48 MFMAs.
// ----------------------------------------------------------------------------
constexpr int ZM_TX = 16, ZM_RW = 8, ZM_TY = 4 * ZM_RW, ZM_HX = ZM_TX + 2, ZM_HY = ZM_TY + 2;
constexpr int ZM_SITES = ZM_HX * ZM_HY;          // 612 per plane
constexpr int ZM_PAD = 40;                       // bf16 per staged site
constexpr int ZM_CHUNKS = ZM_SITES * 4;          // 16-byte chunks per plane

// NW waves per workgroup: 4 (one per SIMD; a wave owns 8 rows and both
// 16-channel blocks, 54 weight fragments in registers) or 8 (two per SIMD,
// which the MFMA pipe needs to approach its peak; wave w owns rows 8 (w & 3)
// .. +7 and channel block w >> 2, 27 weight fragments).  The same products
// in the same order either way.
template <int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4))) void k_conv_z32(
    const bf16 *__restrict__ X, int ldx, int xoff, const bf16 *__restrict__ Wt,
    const float *__restrict__ bias, bf16 *__restrict__ Y, int ldy, int yoff, int B, int D, int H,
    int Wd, int zc, int ntx, int nty)
{
    constexpr int NT = 64 * NW, NJW = NW == 8 ? 1 : 2;
    constexpr int PER_THREAD = (ZM_CHUNKS + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) bf16 zsh[3 * ZM_SITES * ZM_PAD];  // 146,880 bytes
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rg = wave & 3, jb = NW == 8 ? (wave >> 2) : 0;  // row group, first channel block
    int r = blockIdx.x;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    r /= nty;
    const int nzc = (D + zc - 1) / zc;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = tx * ZM_TX, y0 = ty * ZM_TY, z0 = zci * zc, zend = min(z0 + zc, D);
    const int xl = lane & 15, q = lane >> 4;
    // weights: w[tap][j] = Wt[tap][16 (jb + j) + xl][8q .. 8q+7]
    bf16x8 w[27][NJW];
#pragma unroll
    for (int tap = 0; tap < 27; tap++)
#pragma unroll
        for (int j = 0; j < NJW; j++)
            w[tap][j] = *(const bf16x8 *)(Wt + ((size_t)tap * 32 + 16 * (jb + j) + xl) * 32 + 8 * q);
    // this thread's staging chunks: site s = c / 4 (row-major 34 x 18), part c % 4
    auto plane_src = [&](int i, int zz) -> const bf16 * {
        const int c = tid + NT * i, s = c >> 2, part = c & 3;
        const int hy = s / ZM_HX, hx = s - hy * ZM_HX;
        const size_t v = (((size_t)b * D + wrapm(zz, D)) * H + wrapm(y0 + hy - 1, H)) * Wd +
                         wrapm(x0 + hx - 1, Wd);
        return X + v * ldx + xoff + 8 * part;
    };
    // bias of the lane's output channels 16 (jb + j) + 4q .. +3: the
    // accumulators start there
    f32x4 bj[NJW];
#pragma unroll
    for (int j = 0; j < NJW; j++)
        bj[j] = bias4(bias, 16 * (jb + j) + 4 * q);
    bf16x8 pf[PER_THREAD];
    auto load_plane = [&](int zz) {
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++)
            if (tid + NT * i < ZM_CHUNKS)
                pf[i] = *(const bf16x8 *)plane_src(i, zz);
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++) {
            const int c = tid + NT * i;
            if (c < ZM_CHUNKS)
                *(bf16x8 *)(zsh + ((size_t)slot * ZM_SITES + (c >> 2)) * ZM_PAD + 8 * (c & 3)) = pf[i];
        }
    };
    // planes z0-1 and z0 staged, z0+1 in registers
    load_plane(z0 - 1);
    store_plane(0);
    load_plane(z0);
    store_plane(1);
    load_plane(z0 + 1);
    for (int z = z0; z < zend; z++) {
        const int s_m = (z - z0) % 3, s_c = (z - z0 + 1) % 3, s_p = (z - z0 + 2) % 3;
        __syncthreads();  // the slot of plane z-2 is no longer read
        store_plane(s_p);  // plane z+1
        __syncthreads();
        if (z + 1 < zend)
            load_plane(z + 2);  // in flight during this plane's MFMAs
        f32x4 acc[ZM_RW][NJW];
#pragma unroll
        for (int rr = 0; rr < ZM_RW; rr++)
#pragma unroll
            for (int j = 0; j < NJW; j++)
                acc[rr][j] = bj[j];
        const int slots[3] = {s_m, s_c, s_p};
#pragma unroll
        for (int dz = 0; dz < 3; dz++)
#pragma unroll
            for (int dx = 0; dx < 3; dx++) {
                bf16x8 a[ZM_RW + 2];
                const bf16 *base = zsh + ((size_t)slots[dz] * ZM_SITES + (ZM_RW * rg) * ZM_HX + xl + dx) *
                                             ZM_PAD + 8 * q;
#pragma unroll
                for (int rr = 0; rr < ZM_RW + 2; rr++)
                    a[rr] = *(const bf16x8 *)(base + rr * ZM_HX * ZM_PAD);
#pragma unroll
                for (int dy = 0; dy < 3; dy++)
#pragma unroll
                    for (int rr = 0; rr < ZM_RW; rr++)
#pragma unroll
                        for (int j = 0; j < NJW; j++)
                            acc[rr][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                w[dz * 9 + dy * 3 + dx][j], a[rr + dy], acc[rr][j], 0, 0, 0);
            }
        // lane holds channels 16 (jb + j) + 4q .. +3 of site (x0+xl, y0+8 rg+rr)
        const int x = x0 + xl;
        if (x < Wd) {
#pragma unroll
            for (int rr = 0; rr < ZM_RW; rr++) {
                const int y = y0 + ZM_RW * rg + rr;
                if (y >= H)
                    break;
                const size_t o = ((((size_t)b * D + z) * H + y) * Wd + x) * ldy + yoff + 16 * jb + 4 * q;
#pragma unroll
                for (int j = 0; j < NJW; j++)
                    *(bf16x4 *)(Y + o + 16 * j) = leaky4(acc[rr][j]);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Input-plane-stationary z march for Cin = Cout = 32 (the level-1 layers),
// forward (MODE 0).  A workgroup owns a 16 (x) x 4R (y) tile and marches a
// chunk of z planes; wave w computes rows R (w & 3) .. +R-1 and the channel
// block w >> 2, with its 27 weight fragments in registers (two waves per
// SIMD).  Each staged input plane p is read from LDS once and feeds the three
// output planes that need it: p + 1 (taps dz = -1), p (dz = 0) and p - 1
// (dz = +1).  Their accumulators (three sets of R rows) rotate with the
// plane; output p - 1 is complete after plane p.  Per input plane a wave runs
// 3 (dx) x (R + 2) LDS row reads and 27 R MFMAs, 9 R / (R + 2) MFMAs per read
// where k_conv_z32 runs 3 R / (R + 2).  Each accumulator still receives its
// products in the order (dz, dx, dy) from the bias, so the sums are bitwise
// those of k_conv_z32 and k_conv_rows.  Two LDS plane buffers, sites padded
// to 48 bf16 (96 bytes: conflict-free fragment reads), one barrier per plane.
// The host picks z chunks of 3k + 1 planes (k >= 1) that tile D: the zc + 2
// input planes of a chunk are whole triples, and the unrolled triple rotates
// the accumulator sets without copies.  In-plane offsets in 32 bits (H W ldx
// < 2^31).
//
// The epilogue policy EPI stores bf16(leaky(acc)) (EpiStore), stores it and
// its 2x2x2 average pooling (EpiPool, FULL tiles; k_pool_fwd's sums), or
// (inference of the last layer, FULL tiles) applies the network's 1x1x1 head
// in place of k_head_fwd / k_head_ctx: the output plane's bf16 activations go
// to an LDS plane (double-buffered), and after the next plane's barrier each
// wave runs head_mfma on two of its rows, so the outputs are bitwise those
// of k_head_fwd, without the activation tensor's round trip through HBM.
// ----------------------------------------------------------------------------
constexpr int ZS_PAD = LDS_PAD32, ZS_TX = 16, ZS_HX = ZS_TX + 2, ZS_R = 4;

template <int V>
using IC = std::integral_constant<int, V>;

// y = x + head(F) into a packed [V][3] output (k_head_fwd).  get() reads a
// site's input (a plane ahead of its use), put() writes its output.
struct HeadPacked {
    typedef float T;
    const bf16 *whd;
    const float *hb, *x;
    float *out;
    __device__ void get(size_t v, T (&in)[3]) const
    {
        in[0] = x[v * 3 + 0];
        in[1] = x[v * 3 + 1];
        in[2] = x[v * 3 + 2];
    }
    __device__ void put(size_t v, const T (&in)[3], float a0, float a1, float a2) const
    {
        out[v * 3 + 0] = in[0] + a0;
        out[v * 3 + 1] = in[1] + a1;
        out[v * 3 + 2] = in[2] + a2;
    }
};
// a context's fields += head(F) (k_head_ctx)
template <typename TF>
struct HeadCtx {
    typedef TF T;
    const bf16 *whd;
    const float *hb;
    T *s, *l, *c;
    __device__ void get(size_t v, T (&in)[3]) const
    {
        in[0] = s[v];
        in[1] = l[v];
        in[2] = c[v];
    }
    __device__ void put(size_t v, const T (&in)[3], float a0, float a1, float a2) const
    {
        s[v] = (T)((float)in[0] + a0);
        l[v] = (T)((float)in[1] + a1);
        c[v] = (T)((float)in[2] + a2);
    }
};
struct EpiStore {
    typedef float T;
};
// store the output and its 2x2x2 average pooling P [B][D/2][H/2][W/2][32]
// (k_pool_fwd)
struct EpiPool {
    typedef float T;
    bf16 *P;
};

template <int R, bool FULL, class EPI>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_zs32(
    const bf16 *__restrict__ X, int ldx, int xoff, const bf16 *__restrict__ Wt,
    const float *__restrict__ bias, bf16 *__restrict__ Y, int ldy, int yoff, int B, int D, int H,
    int Wd, int zc, int ntx, int nty, const EPI epi)
{
    constexpr bool POOL = std::is_same<EPI, EpiPool>::value;
    constexpr bool HEAD = !std::is_same<EPI, EpiStore>::value && !POOL;
    static_assert(!POOL || (FULL && R % 2 == 0), "pooling runs on full tiles of row pairs");
    constexpr int TY = 4 * R, HY = TY + 2, SITES = ZS_HX * HY, CHUNKS = 4 * SITES;
    constexpr int PT = (CHUNKS + 511) / 512;  // staged 16-byte chunks per thread
    constexpr int HSITES = HEAD ? ZS_TX * TY : 1;
    __shared__ __attribute__((aligned(16))) bf16 zsh[2 * SITES * ZS_PAD];  // 62,208 bytes at R = 4
    __shared__ __attribute__((aligned(16))) bf16 hsh[2 * HSITES * ZS_PAD];  // 49,152 bytes (head)
    __shared__ __attribute__((aligned(16))) bf16 whs[HEAD ? 2 * 16 * 32 : 8];  // head weights (Whd)
    __shared__ float hbs[4];                                                     // head bias
    static_assert(!HEAD || FULL, "the head epilogue runs on full tiles");
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int rg = wave & 3, jb = wave >> 2;
    int r = blockIdx.x;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    r /= nty;
    const int nzc = D / zc;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = tx * ZS_TX, y0 = ty * TY, z0 = zci * zc, zend = z0 + zc;
    const int xl = lane & 15, q = lane >> 4;
    // weights: w[tap] = Wt[tap][16 jb + xl][8q .. 8q+7]
    bf16x8 w[27];
#pragma unroll
    for (int tap = 0; tap < 27; tap++)
        w[tap] = *(const bf16x8 *)(Wt + ((size_t)tap * 32 + 16 * jb + xl) * 32 + 8 * q);
    const f32x4 bias0 = bias4(bias, 16 * jb + 4 * q);
    // staging: chunk c -> site c / 4 (row-major HY x 18), part c % 4; every
    // thread issues the same loads (chunks past the end re-read the last one)
    unsigned fo[PT];
#pragma unroll
    for (int i = 0; i < PT; i++) {
        const int c = min(tid + 512 * i, CHUNKS - 1);
        const int s = c >> 2, part = c & 3;
        const int hy = s / ZS_HX, hx = s - hy * ZS_HX;
        fo[i] = (unsigned)((wrapm(y0 + hy - 1, H) * Wd + wrapm(x0 + hx - 1, Wd)) * ldx + xoff + 8 * part);
    }
    bf16x8 pf[PT];
    auto load_plane = [&](int zz) {
        const bf16 *pl = X + ((size_t)b * D + wrapm(zz, D)) * H * (size_t)Wd * ldx;
#pragma unroll
        for (int i = 0; i < PT; i++)
            pf[i] = *(const bf16x8 *)(pl + fo[i]);
    };
    // (the chunks past the end rewrite the last one with its own value: no
    // branch, so that the compiler can count every thread's memory
    // operations in its waits)
    auto store_plane = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int c = min(tid + 512 * i, CHUNKS - 1);
            *(bf16x8 *)(zsh + ((size_t)buf * SITES + (c >> 2)) * ZS_PAD + 8 * (c & 3)) = pf[i];
        }
    };
    // lane holds channels 16 jb + 4q .. +3 of site (x0 + xl, y0 + R rg + rr).
    // FULL (the tiles cover the plane exactly): every store is issued, so
    // the waits before the next plane's staging leave them in flight; vmcnt
    // counts stores too, and a store the compiler cannot count makes it wait
    // for all of them.
    // pooling (EpiPool): the even lane of an x pair sums, for row pair j and
    // its four channels, the pair's values in k_pool_fwd's order (from 0:
    // (dz, dy, dx) = 000, 001, 010, 011 on the even plane, kept in pp, then
    // 100 .. 111 on the odd plane) and stores the average
    float pp[R / 2][4];
    auto pool_out = [&](const bf16x4(&o)[R], int zz) {
        bf16x4 nb[R];  // the x + 1 neighbour's values (lane xl ^ 1)
#pragma unroll
        for (int rr = 0; rr < R; rr++) {
            const uint2 u = __builtin_bit_cast(uint2, o[rr]);
            uint2 v;
            v.x = __builtin_amdgcn_mov_dpp(u.x, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
            v.y = __builtin_amdgcn_mov_dpp(u.y, 0xB1, 0xF, 0xF, false);
            nb[rr] = __builtin_bit_cast(bf16x4, v);
        }
        const bool odd = zz & 1;
#pragma unroll
        for (int j = 0; j < R / 2; j++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                float t = odd ? pp[j][e] : 0.f;
                t += (float)o[2 * j][e];
                t += (float)nb[2 * j][e];
                t += (float)o[2 * j + 1][e];
                t += (float)nb[2 * j + 1][e];
                pp[j][e] = t;
            }
        if constexpr (POOL)
            if (odd && !(xl & 1)) {
            const int H2 = H >> 1, W2 = Wd >> 1;
#pragma unroll
            for (int j = 0; j < R / 2; j++) {
                bf16x4 r4;
#pragma unroll
                for (int e = 0; e < 4; e++)
                    r4[e] = (bf16)(pp[j][e] * 0.125f);
                const size_t pv = (((size_t)b * (D >> 1) + (zz >> 1)) * H2 + ((y0 + R * rg) >> 1) + j) * W2 +
                                  ((x0 + xl) >> 1);
                *(bf16x4 *)(epi.P + pv * 32 + 16 * jb + 4 * q) = r4;
            }
        }
    };
    auto store_out = [&](const f32x4(&S)[R], int zz, int k) {
        if (HEAD) {
            bf16 *hb = hsh + (size_t)(k & 1) * HSITES * ZS_PAD + 16 * jb + 4 * q;
#pragma unroll
            for (int rr = 0; rr < R; rr++)
                *(bf16x4 *)(hb + ((R * rg + rr) * ZS_TX + xl) * ZS_PAD) = leaky4(S[rr]);
            return;
        }
        const int x = x0 + xl;
        if constexpr (POOL) {
            bf16x4 o4[R];
#pragma unroll
            for (int rr = 0; rr < R; rr++) {
                const int y = y0 + R * rg + rr;
                const size_t o = ((((size_t)b * D + zz) * H + y) * Wd + x) * ldy + yoff + 16 * jb + 4 * q;
                o4[rr] = leaky4(S[rr]);
                *(bf16x4 *)(Y + o) = o4[rr];
            }
            pool_out(o4, zz);
        } else if (FULL || x < Wd) {
#pragma unroll
            for (int rr = 0; rr < R; rr++) {
                const int y = y0 + R * rg + rr;
                if (!FULL && y >= H)
                    break;
                const size_t o = ((((size_t)b * D + zz) * H + y) * Wd + x) * ldy + yoff + 16 * jb + 4 * q;
                *(bf16x4 *)(Y + o) = leaky4(S[rr]);
            }
        }
    };
    // head pass (HEAD): the output plane's 16 rows of 16 sites, two rows per
    // wave (wave w: rows 2w, 2w + 1), each by head_mfma on the LDS plane; the
    // lanes of q = 0 read their two sites' inputs a plane ahead (head_load)
    // and write the outputs
    static_assert(!HEAD || TY == 16, "the head pass covers 16 rows with 8 waves");
    typename EPI::T hin[2][3];
    auto head_site = [&](int zz, int t) {
        return (((size_t)b * D + zz) * H + y0 + 2 * wave + t) * Wd + x0 + xl;
    };
    auto head_load = [&](int zz) {
        if constexpr (HEAD) {
            if (q == 0)
#pragma unroll
                for (int t = 0; t < 2; t++)
                    epi.get(head_site(zz, t), hin[t]);
        }
    };
    auto head_pass = [&](int zz, int hb_k) {
        if constexpr (HEAD) {
            const bf16 *hp = hsh + ((size_t)(hb_k & 1) * HSITES + 2 * wave * ZS_TX + xl) * ZS_PAD + 8 * q;
            f32x4 o[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const bf16x8 f[1] = {*(const bf16x8 *)(hp + t * ZS_TX * ZS_PAD)};
                o[t] = head_mfma<1>(whs, hbs, f, lane);
            }
            if (q == 0)
#pragma unroll
                for (int t = 0; t < 2; t++)
                    epi.put(head_site(zz, t), hin[t], o[t][0], o[t][1], o[t][2]);
        }
    };
    const int lofs = ((R * rg) * ZS_HX + xl) * ZS_PAD + 8 * q;
    // One input plane (k: its index in the chunk, buffer k & 1).  MSK bit 2:
    // output p + 1 (Sp), bit 1: output p (S0), bit 0: output p - 1 (Sm).
    // The 3 (R + 2) row reads (dx, row) run PD ahead of their MFMAs in a
    // ring of PD + 1 (one ahead with the head epilogue, whose registers it
    // needs; the distance made no measurable difference).
    constexpr int PD = HEAD ? 1 : 2;
    auto plane = [&](auto MSK, f32x4(&Sp)[R], f32x4(&S0)[R], f32x4(&Sm)[R], int k) {
        constexpr int M = decltype(MSK)::value, NR = 3 * (R + 2);
        const bf16 *pb = zsh + (size_t)(k & 1) * SITES * ZS_PAD + lofs;
        auto rd = [&](int idx) {
            const int dx = idx / (R + 2), rr = idx % (R + 2);
            return *(const bf16x8 *)(pb + (rr * ZS_HX + dx) * ZS_PAD);
        };
        bf16x8 a[PD + 1];
#pragma unroll
        for (int i = 0; i < PD; i++)
            a[i] = rd(i);
#pragma unroll
        for (int idx = 0; idx < NR; idx++) {
            if (idx + PD < NR)
                a[(idx + PD) % (PD + 1)] = rd(idx + PD);
            __builtin_amdgcn_sched_barrier(0);
            const int dx = idx / (R + 2), rr = idx % (R + 2);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int ro = rr - dy;
                if (ro >= 0 && ro < R) {
                    const bf16x8 av = a[idx % (PD + 1)];
                    if (M & 4)
                        Sp[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[dy * 3 + dx], av, Sp[ro], 0, 0, 0);
                    if (M & 2)
                        S0[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[9 + dy * 3 + dx], av, S0[ro], 0, 0, 0);
                    if (M & 1)
                        Sm[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[18 + dy * 3 + dx], av, Sm[ro], 0, 0,
                                                                         0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // Plane k of the chunk (p = z0 - 1 + k): stage plane p + 1 (STG), run
    // the head of output p - 2 (HP) and load the inputs of output p - 1's,
    // start loading plane p + 2 (LDG), run plane p, finish output p - 1
    // (MSK bit 0) and restart its accumulators at the bias for output p + 2.
    auto step = [&](auto MSK, auto STG, auto LDG, auto HP, f32x4(&Sp)[R], f32x4(&S0)[R],
                    f32x4(&Sm)[R], int k) {
        const int p = z0 - 1 + k;
        __syncthreads();  // plane p is staged; plane p - 1's buffer is no longer read
        if (decltype(STG)::value)
            store_plane((k + 1) & 1);
        if (decltype(HP)::value)  // (between the staging and the next loads: pf is free)
            head_pass(p - 2, k - 1);
        if (decltype(HP)::value || k == 2)
            head_load(p - 1);  // the next head pass's inputs, a plane ahead
        if (decltype(LDG)::value)
            load_plane(p + 2);
        plane(MSK, Sp, S0, Sm, k);
        if (decltype(MSK)::value & 1)
            store_out(Sm, p - 1, k);
#pragma unroll
        for (int rr = 0; rr < R; rr++)
            Sm[rr] = bias0;
    };
    f32x4 A0[R], A1[R], A2[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++)
        A0[rr] = A1[rr] = A2[rr] = bias0;
    if constexpr (HEAD) {
        // the head weights and bias go to LDS once (published by the first
        // plane's barrier): a global load in the head pass would wait for the
        // plane loads and stores in flight (vmcnt counts them all)
        if (tid < 2 * 16 * 32 / 8)
            *(bf16x8 *)(whs + 8 * tid) = *(const bf16x8 *)(epi.whd + 8 * tid);
        if (tid < 3)
            hbs[tid] = epi.hb[tid];
    }
#pragma unroll
    for (int i = 0; i < 27; i++)
        settle(w[i]);
    settle(bias0);
    load_plane(z0 - 1);
    store_plane(0);
    load_plane(z0);
    // triple t: planes k = 3t, 3t + 1, 3t + 2 with (Sp, S0, Sm) = (A0, A2, A1),
    // (A1, A0, A2), (A2, A1, A0); the first triple has no output z0 - 1 or
    // z0 - 2, the last none past zend - 1
    const int T = (zc + 2) / 3;
    step(IC<4>{}, IC<1>{}, IC<1>{}, IC<0>{}, A0, A2, A1, 0);
    step(IC<6>{}, IC<1>{}, IC<1>{}, IC<0>{}, A1, A0, A2, 1);
    step(IC<7>{}, IC<1>{}, IC<1>{}, IC<0>{}, A2, A1, A0, 2);
    for (int t = 1; t < T - 1; t++) {
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<HEAD>{}, A0, A2, A1, 3 * t);
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<HEAD>{}, A1, A0, A2, 3 * t + 1);
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<HEAD>{}, A2, A1, A0, 3 * t + 2);
    }
    step(IC<7>{}, IC<1>{}, IC<1>{}, IC<HEAD>{}, A0, A2, A1, 3 * (T - 1));
    step(IC<3>{}, IC<1>{}, IC<0>{}, IC<HEAD>{}, A1, A0, A2, 3 * (T - 1) + 1);
    step(IC<1>{}, IC<0>{}, IC<0>{}, IC<HEAD>{}, A2, A1, A0, 3 * (T - 1) + 2);
    if (HEAD) {
        __syncthreads();
        head_pass(zend - 1, 3 * (T - 1) + 2);
    }
}

// ----------------------------------------------------------------------------
// L8 of the base-32 network (Cout 32, 32 skip and 64 upsampled channels):
// k_conv_poly's products (upsampled channels in polyphase form, the 27 taps
// of the skip channels), z-marching with both sources staged in LDS.  A
// workgroup owns a fine 32 (x) x 8 (y) tile and marches a chunk of z planes.
// LDS holds three fine planes of the skip channels (tile + one-site halo,
// 10 rows x 34 sites, stored by x parity so that a wave's lanes read
// consecutive sites; 32 channels padded to 48) and two coarse planes of the
// upsampled source (6 rows x 18 sites, 64 channels padded to 72).  A fine
// plane z reads coarse planes (z >> 1) + (z & 1) - 1 and the next one, so a
// new coarse plane is staged every second fine plane, over the one that the
// plane being computed no longer reads.  D, H, W even; chunks start at even
// z.
// Each output (x parity px: sites x0 + 2 l + px; 16 channels, nt) is shared
// by two waves that split the K dimension, so that every SIMD runs two waves
// of eight accumulator chains (one wave per SIMD with eight chains issues
// MFMAs at about a third of the peak rate; two waves, above four fifths):
//   waves 0-3 ("up"): the 8 upsampled-channel groups ((chunk, oz, ox), 16
//     MFMAs each, summed weights streamed from the cache four groups ahead,
//     one contiguous kilobyte per fragment), then the skip groups (dz, dx) =
//     (2, 0), (2, 1): 176 MFMAs per plane; the partial sum (from zero) goes
//     to LDS;
//   waves 4-7 ("skip"): the skip groups (0, 0..2), (1, 0..2), (2, 2), weights
//     in registers: 168 MFMAs per plane, from the bias; after the next
//     plane's first barrier they add the partial and store.
// The sum is k_conv_poly's products in another association (two partial
// sums), so the outputs agree with it to FP32 rounding, not bitwise.  LDS:
// 97,920 + 31,104 bytes of planes and 32,768 of partial sums.
// ----------------------------------------------------------------------------
constexpr int ZP_TX = 32, ZP_TY = 8;
constexpr int ZP_FX = 17, ZP_FROWS = ZP_TY + 2, ZP_FPAD = LDS_PAD32;
constexpr int ZP_FSITES = ZP_FROWS * 2 * ZP_FX;  // 340 per fine plane
constexpr int ZP_CX = ZP_TX / 2 + 2, ZP_CROWS = ZP_TY / 2 + 2;
constexpr int ZP_CSITES = ZP_CROWS * ZP_CX;      // 108 per coarse plane
constexpr int ZP_FCH = ZP_FSITES * 4;            // 16-byte chunks per plane
constexpr int ZP_CCH = ZP_CSITES * 8;
constexpr int ZQ_CPAD = 72;
constexpr int ZQ_FPT = (ZP_FCH + 511) / 512, ZQ_CPT = (ZP_CCH + 511) / 512;

// FULL (the tiles cover the grid): the output and staging stores are
// issued unconditionally, so that the compiler can count them: vmcnt counts
// stores too, and a store it cannot count makes the skip waves' wait for the
// next planes' staging loads a wait for their output stores to complete.
template <bool FULL>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_zpoly8(
    const bf16 *__restrict__ X, int ldx, int xoff, const bf16 *__restrict__ Wt,
    const bf16 *__restrict__ Wp, const float *__restrict__ bias, bf16 *__restrict__ Y, int ldy,
    int B, int D, int H, int Wd, const bf16 *__restrict__ Xc, int ldc, int zc, int ntx, int nty)
{
    __shared__ __attribute__((aligned(16))) bf16 fsh[3 * ZP_FSITES * ZP_FPAD];  // 97,920 bytes
    __shared__ __attribute__((aligned(16))) bf16 csh[2 * ZP_CSITES * ZQ_CPAD];  // 31,104 bytes
    __shared__ f32x4 red[4 * ZP_TY * 64];                                       // 32,768 bytes
    constexpr int CIN = 96, CUP = 64, COUT = 32;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int px = wave & 1, nt = (wave >> 1) & 1, role = wave >> 2;
    int r = blockIdx.x;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    r /= nty;
    const int nzc = (D + zc - 1) / zc;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = tx * ZP_TX, y0 = ty * ZP_TY, z0 = zci * zc, zend = min(z0 + zc, D);
    const int Dc = D / 2, Hc = H / 2, Wc = Wd / 2, Z0 = z0 / 2;
    const int xl = lane & 15, q = lane >> 4;
    // staging: fine chunk c -> site c / 4 (LDS order: row, parity, index),
    // part c % 4; coarse chunk c -> site c / 8, part c % 8.  Every thread
    // issues the same loads (chunks past the end re-read the last one and
    // are not stored), so the compiler can count them in its waits; each
    // thread's in-plane element offsets (planes of under 2^31 elements) are
    // computed once
    bf16x8 pf[ZQ_FPT], pc[ZQ_CPT];
    unsigned fo[ZQ_FPT], co[ZQ_CPT];
#pragma unroll
    for (int i = 0; i < ZQ_FPT; i++) {
        const int c = min(tid + 512 * i, ZP_FCH - 1);
        const int s = c >> 2, part = c & 3;
        const int fr = s / (2 * ZP_FX), rem = s - fr * 2 * ZP_FX;
        const int par = rem / ZP_FX, idx = rem - par * ZP_FX;
        const int lx = 2 * idx + par;
        fo[i] = (unsigned)((wrapm(y0 + fr - 1, H) * Wd + wrapm(x0 + lx - 1, Wd)) * ldx + xoff + 8 * part);
    }
#pragma unroll
    for (int i = 0; i < ZQ_CPT; i++) {
        const int c = min(tid + 512 * i, ZP_CCH - 1);
        const int s = c >> 3, part = c & 7;
        const int cr = s / ZP_CX, cx = s - cr * ZP_CX;
        co[i] = (unsigned)((wrapm(y0 / 2 + cr - 1, Hc) * Wc + wrapm(x0 / 2 + cx - 1, Wc)) * ldc + 8 * part);
    }
    auto load_fine = [&](int zz) {
        const bf16 *pl = X + ((size_t)b * D + wrapm(zz, D)) * H * (size_t)Wd * ldx;
#pragma unroll
        for (int i = 0; i < ZQ_FPT; i++)
            pf[i] = *(const bf16x8 *)(pl + fo[i]);
    };
    // (FULL: the chunks past the end rewrite the last one with its own value)
    auto store_fine = [&](int slot) {
#pragma unroll
        for (int i = 0; i < ZQ_FPT; i++) {
            const int c = FULL ? min(tid + 512 * i, ZP_FCH - 1) : tid + 512 * i;
            if (FULL || c < ZP_FCH)
                *(bf16x8 *)(fsh + ((size_t)slot * ZP_FSITES + (c >> 2)) * ZP_FPAD + 8 * (c & 3)) = pf[i];
        }
    };
    auto load_coarse = [&](int zz) {
        const bf16 *pl = Xc + ((size_t)b * Dc + wrapm(zz, Dc)) * Hc * (size_t)Wc * ldc;
#pragma unroll
        for (int i = 0; i < ZQ_CPT; i++)
            pc[i] = *(const bf16x8 *)(pl + co[i]);
    };
    auto store_coarse = [&](int slot) {
#pragma unroll
        for (int i = 0; i < ZQ_CPT; i++) {
            const int c = FULL ? min(tid + 512 * i, ZP_CCH - 1) : tid + 512 * i;
            if (FULL || c < ZP_CCH)
                *(bf16x8 *)(csh + ((size_t)slot * ZP_CSITES + (c >> 3)) * ZQ_CPAD + 8 * (c & 7)) = pc[i];
        }
    };
    // fine plane P in slot (P - z0 + 1) % 3, coarse plane P in (P - Z0 + 1) & 1
    load_fine(z0 - 1);
    store_fine(0);
    load_fine(z0);
    store_fine(1);
    load_coarse(Z0 - 1);
    store_coarse(0);
    load_coarse(Z0);
    store_coarse(1);
    load_fine(z0 + 1);
    // skip-channel weight fragment of tap t: Wt[t][16 nt + xl][64 + 8q ..]
    auto wskip = [&](int t) {
        return *(const bf16x8 *)(Wt + ((size_t)t * COUT + 16 * nt + xl) * CIN + CUP + 8 * q);
    };
    // fragment reads of one group: fine rows of plane slot fs at offset dx,
    // or coarse rows of slot cs at offset ox, chunk kc
    auto ld_fine = [&](bf16x8(&buf)[ZP_FROWS], int fs, int dx) {
        const int par = (px + dx) & 1, idx = xl + ((px + dx) >> 1);
        const bf16 *fb = fsh + ((size_t)fs * ZP_FSITES + par * ZP_FX + idx) * ZP_FPAD + 8 * q;
#pragma unroll
        for (int rr = 0; rr < ZP_FROWS; rr++)
            buf[rr] = *(const bf16x8 *)(fb + rr * 2 * ZP_FX * ZP_FPAD);
    };
    auto ld_coarse = [&](bf16x8(&buf)[ZP_FROWS], int cs, int ox, int kc) {
        const bf16 *cb = csh + ((size_t)cs * ZP_CSITES + xl + px + ox) * ZQ_CPAD + kc + 8 * q;
#pragma unroll
        for (int cr = 0; cr < ZP_CROWS; cr++)
            buf[cr] = *(const bf16x8 *)(cb + cr * ZP_CX * ZQ_CPAD);
    };
    f32x4 *const redp = red + (px + 2 * nt) * ZP_TY * 64 + lane;
    f32x4 acc[ZP_TY];
    if (role == 0) {
        // ---- up waves ----
        bf16x8 ws[6];  // taps (dz, dy, dx) = (2, dy, 0), (2, dy, 1)
#pragma unroll
        for (int dx = 0; dx < 2; dx++)
#pragma unroll
            for (int dy = 0; dy < 3; dy++)
                ws[3 * dx + dy] = wskip(18 + 3 * dy + dx);
        const __amdgpu_buffer_rsrc_t wpr = __builtin_amdgcn_make_buffer_rsrc(
            (void *)Wp, (short)0, 64 * COUT * CUP * 2, 0x00020000);
        const int wpo = ((16 * nt + xl) * 32 + 8 * q) * 2;
        // summed weights of group u (chunk u >> 2, oz = (u >> 1) & 1, ox = u & 1)
        // for plane parity pzz, ring slot u & 3 (layout: k_prep_wpoly)
        bf16x8 wq[4][2][2];
        auto load_wq = [&](int pzz, int u) {
#pragma unroll
            for (int py = 0; py < 2; py++)
#pragma unroll
                for (int oy = 0; oy < 2; oy++)
                    wq[u & 3][py][oy] = __builtin_bit_cast(
                        bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                    wpr, wpo,
                                    ((((4 * pzz + 2 * py + px) * 8 + 4 * ((u >> 1) & 1) + 2 * oy + (u & 1)) *
                                          (CUP / 32) + (u >> 2)) * COUT * 32) * 2,
                                    0));
        };
#pragma unroll
        for (int u = 0; u < 4; u++)
            load_wq(z0 & 1, u);
#pragma unroll
        for (int i = 0; i < 6; i++)
            settle(ws[i]);
        for (int z = z0; z < zend; z++) {
            const int pz = z & 1;
            __syncthreads();  // every read of the previous plane is done
            store_fine((z - z0 + 2) % 3);
            if (pz)
                store_coarse(((z >> 1) + 1 - Z0 + 1) & 1);
            __syncthreads();
            // (this wave's share of the next planes is loaded after its last
            // weight load: a use waits for every load issued before it)
#pragma unroll
            for (int rr = 0; rr < ZP_TY; rr++)
                acc[rr] = f32x4{0.f, 0.f, 0.f, 0.f};
            auto ldg = [&](bf16x8(&buf)[ZP_FROWS], int g) {
                if (g < 8) {
                    const int oz = (g >> 1) & 1;
                    ld_coarse(buf, ((z >> 1) + pz - 1 + oz - Z0 + 1) & 1, g & 1, 32 * (g >> 2));
                } else {
                    ld_fine(buf, (z - z0 + 2) % 3, g - 8);
                }
            };
            auto mmg = [&](const bf16x8(&buf)[ZP_FROWS], int g) {
                if (g < 8) {
#pragma unroll
                    for (int oy = 0; oy < 2; oy++)
#pragma unroll
                        for (int rr = 0; rr < ZP_TY; rr++)
                            acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                wq[g & 3][rr & 1][oy], buf[(rr >> 1) + (rr & 1) + oy], acc[rr], 0, 0, 0);
                } else {
#pragma unroll
                    for (int dy = 0; dy < 3; dy++)
#pragma unroll
                        for (int rr = 0; rr < ZP_TY; rr++)
                            acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                ws[3 * (g - 8) + dy], buf[rr + dy], acc[rr], 0, 0, 0);
                }
            };
            bf16x8 b0[ZP_FROWS], b1[ZP_FROWS];
            ldg(b0, 0);
#pragma unroll
            for (int g = 0; g < 10; g++) {
                if (g % 2 == 0) {
                    if (g + 1 < 10)
                        ldg(b1, g + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    mmg(b0, g);
                } else {
                    if (g + 1 < 10)
                        ldg(b0, g + 1);
                    __builtin_amdgcn_sched_barrier(0);
                    mmg(b1, g);
                }
                __builtin_amdgcn_sched_barrier(0);
                if (g < 4)
                    load_wq(pz, g + 4);  // into the slot group g has just read
                if (g == 3) {
                    load_fine(z + 2);
                    load_coarse(((z + 1) >> 1) + 1);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int rr = 0; rr < ZP_TY; rr++)
                redp[rr * 64] = acc[rr];
#pragma unroll
            for (int u = 0; u < 4; u++)
                load_wq(pz ^ 1, u);  // the next plane's (read and not used after the last)
        }
        __syncthreads();
    } else {
        // ---- skip waves ----
        bf16x8 ws[21];  // groups (dz, dx): 0..5 = (dz 0..1, dx 0..2), 6 = (2, 2)
#pragma unroll
        for (int g = 0; g < 7; g++)
#pragma unroll
            for (int dy = 0; dy < 3; dy++)
                ws[3 * g + dy] = wskip(g < 6 ? (g / 3) * 9 + 3 * dy + g % 3 : 18 + 3 * dy + 2);
        const f32x4 bias0 = bias4(bias, 16 * nt + 4 * q);
        // lane holds channels 16 nt + 4q .. +3 of site (x0 + 2 xl + px, y0 + rr)
        auto add_partial = [&]() {
#pragma unroll
            for (int rr = 0; rr < ZP_TY; rr++)
                acc[rr] += redp[rr * 64];
        };
        auto store_out = [&](int zz) {
            const int x = x0 + 2 * xl + px;
            if (FULL || x < Wd) {
#pragma unroll
                for (int rr = 0; rr < ZP_TY; rr++) {
                    const int y = y0 + rr;
                    if (!FULL && y >= H)
                        break;
                    const size_t o = ((((size_t)b * D + zz) * H + y) * Wd + x) * ldy + 16 * nt + 4 * q;
                    *(bf16x4 *)(Y + o) = leaky4(acc[rr]);
                }
            }
        };
        // the weights and bias have arrived before the march (settle)
#pragma unroll
        for (int i = 0; i < 21; i++)
            settle(ws[i]);
        settle(bias0);
        // one plane; FIRST (the chunk's first plane: no output of the plane
        // before it) and ODD are static, so that the output stores are
        // unconditional in the steady state and the compiler counts them in
        // its waits for the staging loads (z0 and the chunk length are even)
        auto plane = [&](auto FIRST, auto ODD, int z) {
            __syncthreads();  // every read of the previous plane is done, its partial written
            if (!decltype(FIRST)::value) {
                add_partial();
                store_out(z - 1);
            }
            store_fine((z - z0 + 2) % 3);
            if (decltype(ODD)::value)
                store_coarse(((z >> 1) + 1 - Z0 + 1) & 1);
            __syncthreads();
            load_fine(z + 2);
            load_coarse(((z + 1) >> 1) + 1);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int rr = 0; rr < ZP_TY; rr++)
                acc[rr] = bias0;
            // one row buffer, walked by anti-diagonals: step i runs the MFMAs
            // that read row i (rows rr = i - dy), after which the row is
            // reloaded with the next group's row i (a group ahead)
            auto row_ptr = [&](int g) {
                const int dz = g < 6 ? g / 3 : 2, dx = g < 6 ? g % 3 : 2;
                const int par = (px + dx) & 1, idx = xl + ((px + dx) >> 1);
                return fsh + ((size_t)((z - z0 + dz) % 3) * ZP_FSITES + par * ZP_FX + idx) * ZP_FPAD + 8 * q;
            };
            bf16x8 buf[ZP_FROWS];
            {
                const bf16 *fb = row_ptr(0);
#pragma unroll
                for (int i = 0; i < ZP_FROWS; i++)
                    buf[i] = *(const bf16x8 *)(fb + i * 2 * ZP_FX * ZP_FPAD);
            }
#pragma unroll
            for (int g = 0; g < 7; g++) {
                const bf16 *fn = row_ptr(g + 1 < 7 ? g + 1 : g);
#pragma unroll
                for (int i = 0; i < ZP_FROWS; i++) {
#pragma unroll
                    for (int dy = 0; dy < 3; dy++) {
                        const int rr = i - dy;
                        if (rr >= 0 && rr < ZP_TY)
                            acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ws[3 * g + dy], buf[i],
                                                                              acc[rr], 0, 0, 0);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                    if (g + 1 < 7)
                        buf[i] = *(const bf16x8 *)(fn + i * 2 * ZP_FX * ZP_FPAD);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        };
        plane(IC<1>{}, IC<0>{}, z0);
        plane(IC<0>{}, IC<1>{}, z0 + 1);
        for (int z = z0 + 2; z < zend; z += 2) {
            plane(IC<0>{}, IC<0>{}, z);
            plane(IC<0>{}, IC<1>{}, z + 1);
        }
        __syncthreads();
        add_partial();
        store_out(zend - 1);
    }
}

// ----------------------------------------------------------------------------
// z-marching convolution for Cin = 32 CK (CK = 1, 2) and Cout = 64, forward
// (MODE 0): the 128^3-level layers of the base-32 network (L2, L3, L7) and
// the level-1 layers of the base-64 one.  A workgroup owns a 16 x 16 tile
// and marches a chunk of z planes; wave w owns output channels 16w .. 16w+15
// of all 16 rows, with its 27 CK weight fragments in registers (one wave per
// SIMD).  Each input plane (tile + one-site halo, 18 x 18 sites, channels
// padded to 48 or 80, LDS_PAD32 / LDS_PAD64) is staged once in LDS for the
// three output planes that need it.  Per output plane and 32-channel chunk a wave runs 9 (dz, dx)
// groups of 18 LDS fragment reads and 48 MFMAs, in k_conv_rows' order
// (chunk, dz, dx, dy), so the sums are bitwise those of k_conv_rows.
// ----------------------------------------------------------------------------
constexpr int Z64_T = 16, Z64_H = Z64_T + 2, Z64_SITES = Z64_H * Z64_H;  // 324 sites

template <int CK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_conv_z64(
    const bf16 *__restrict__ X, int ldx, int xoff, const bf16 *__restrict__ Wt,
    const float *__restrict__ bias, bf16 *__restrict__ Y, int ldy, int yoff, int B, int D, int H,
    int Wd, int zc, int ntx, int nty)
{
    constexpr int CIN = 32 * CK, PAD = CK == 1 ? LDS_PAD32 : LDS_PAD64;
    constexpr int PARTS = CIN / 8;                        // 16-byte chunks per site
    constexpr int CHUNKS = Z64_SITES * PARTS;             // per plane
    constexpr int PER_THREAD = (CHUNKS + 255) / 256;
    __shared__ __attribute__((aligned(16))) bf16 zsh[3 * Z64_SITES * PAD];  // 155,520 bytes at CK 2
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int r = blockIdx.x;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    r /= nty;
    const int nzc = (D + zc - 1) / zc;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = tx * Z64_T, y0 = ty * Z64_T, z0 = zci * zc, zend = min(z0 + zc, D);
    const int xl = lane & 15, q = lane >> 4;
    // weights of this wave's 16 channels: w[tap][k] = Wt[tap][16 wave + xl][32k + 8q ..]
    bf16x8 w[27][CK];
#pragma unroll
    for (int tap = 0; tap < 27; tap++)
#pragma unroll
        for (int k = 0; k < CK; k++)
            w[tap][k] = *(const bf16x8 *)(Wt + ((size_t)tap * 64 + 16 * wave + xl) * CIN + 32 * k + 8 * q);
    auto plane_src = [&](int i, int zz) -> const bf16 * {
        const int c = tid + 256 * i, s = c / PARTS, part = c % PARTS;
        const int hy = s / Z64_H, hx = s - hy * Z64_H;
        const size_t v = (((size_t)b * D + wrapm(zz, D)) * H + wrapm(y0 + hy - 1, H)) * Wd +
                         wrapm(x0 + hx - 1, Wd);
        return X + v * ldx + xoff + 8 * part;
    };
    const f32x4 bias0 = bias4(bias, 16 * wave + 4 * q);
    bf16x8 pf[PER_THREAD];
    auto load_plane = [&](int zz) {
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++)
            if (tid + 256 * i < CHUNKS)
                pf[i] = *(const bf16x8 *)plane_src(i, zz);
    };
    auto store_plane = [&](int slot) {
#pragma unroll
        for (int i = 0; i < PER_THREAD; i++) {
            const int c = tid + 256 * i;
            if (c < CHUNKS)
                *(bf16x8 *)(zsh + ((size_t)slot * Z64_SITES + c / PARTS) * PAD + 8 * (c % PARTS)) = pf[i];
        }
    };
    load_plane(z0 - 1);
    store_plane(0);
    load_plane(z0);
    store_plane(1);
    load_plane(z0 + 1);
    for (int z = z0; z < zend; z++) {
        const int s_m = (z - z0) % 3, s_c = (z - z0 + 1) % 3, s_p = (z - z0 + 2) % 3;
        __syncthreads();  // the slot of plane z-2 is no longer read
        store_plane(s_p);  // plane z+1
        __syncthreads();
        if (z + 1 < zend)
            load_plane(z + 2);  // in flight during this plane's MFMAs
        f32x4 acc[Z64_T];
#pragma unroll
        for (int rr = 0; rr < Z64_T; rr++)
            acc[rr] = bias0;
        const int slots[3] = {s_m, s_c, s_p};
#pragma unroll
        for (int k = 0; k < CK; k++)
#pragma unroll
            for (int dz = 0; dz < 3; dz++)
#pragma unroll
                for (int dx = 0; dx < 3; dx++) {
                    bf16x8 a[Z64_H];
                    const bf16 *base = zsh + ((size_t)slots[dz] * Z64_SITES + xl + dx) * PAD + 32 * k + 8 * q;
#pragma unroll
                    for (int rr = 0; rr < Z64_H; rr++)
                        a[rr] = *(const bf16x8 *)(base + rr * Z64_H * PAD);
#pragma unroll
                    for (int dy = 0; dy < 3; dy++)
#pragma unroll
                        for (int rr = 0; rr < Z64_T; rr++)
                            acc[rr] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                w[dz * 9 + dy * 3 + dx][k], a[rr + dy], acc[rr], 0, 0, 0);
                }
        // lane holds channels 16 wave + 4q .. +3 of site (x0+xl, y0+rr) in acc[rr]
        const int x = x0 + xl;
        if (x < Wd) {
#pragma unroll
            for (int rr = 0; rr < Z64_T; rr++) {
                const int y = y0 + rr;
                if (y >= H)
                    break;
                const size_t o = ((((size_t)b * D + z) * H + y) * Wd + x) * ldy + yoff + 16 * wave + 4 * q;
                *(bf16x4 *)(Y + o) = leaky4(acc[rr]);
            }
        }
    }
}

// ----------------------------------------------------------------------------
// Input-plane-stationary z march for Cin = 32 CK (CK = 1, 2) and Cout = 64,
// forward (MODE 0): the 128^3-level layers of the base-32 network (L2, L3,
// L7).  As k_conv_zs32: each staged input plane is read from LDS once for the
// three output planes that need it (three rotating accumulator sets of 8
// rows), two plane buffers, one barrier per plane, z chunks of 3k + 1
// planes, FULL tiles only.  Eight waves, two per SIMD: wave w owns output
// channels 16 (w & 3) .. +15 with 27 weight fragments in registers, and
// (R = 4 rows per wave; 8 spill)
//   CK = 1: rows R (w >> 2) .. +R-1 of a 16 x 2R tile; the sums are bitwise
//     those of k_conv_z64 and k_conv_rows;
//   CK = 2: input channels 32 (w >> 2) .. +31 of all R rows of a 16 x R
//     tile.  The two waves of an output block sum the two halves of K from
//     the bias and from zero; when an output plane is complete both write
//     their sums to LDS (double-buffered) and, after the next plane's
//     barrier, each adds them for half of the rows (bias + first half, plus
//     second half), applies the activation and stores.  The association differs from k_conv_rows'
//     (one running sum over both halves) by FP32 rounding.
// ----------------------------------------------------------------------------
constexpr int ZS64_R = 4;  // rows per wave (8 spill)

// A launch computes the 64 output channels co0 .. co0 + 63 of a layer whose
// weights are Wt [27][wcout][wld], its input channels starting at woff (L6's
// skip channels: wld 192, woff 128; the 64^3 layers: wcout 128, co0 0 or
// 64, woff 0 or 64).  Bias and outputs are those of channels co0 ..; the
// output element of a site is yoff + its channel offset from co0.
// MODE (CK = 2):
//   0: from the bias, bf16(leaky(sum)) into Y;
//   1 (partial input): from FP32 partial sums P [site][pld] (columns co0
//     ..): both K halves sum from zero and the combine adds (P + first half)
//     + second half, P read a plane ahead; bf16(leaky) into Y;
//   2 (partial output): from the bias, the plain FP32 sum into P;
//   3: as 0, and the 2x2x2 average pooling of the outputs into Q [B][D/2]
//     [H/2][W/2][64] (k_pool_fwd's sums: the even lane of an x pair adds the
//     pair's two rows on the even plane, keeps the sum, adds the odd
//     plane's and stores; z chunks start at even planes).
template <int CK, int MODE = 0>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_conv_zs64(
    const bf16 *__restrict__ X, int ldx, int xoff, const bf16 *__restrict__ Wt,
    const float *__restrict__ bias, bf16 *__restrict__ Y, int ldy, int yoff, int B, int D, int H,
    int Wd, int zc, int ntx, int nty, int wcout, int co0, int wld, int woff, float *__restrict__ P,
    int pld, bf16 *__restrict__ Q = nullptr)
{
    constexpr bool PIN = MODE == 1, POUT = MODE == 2, POOL = MODE == 3;
    static_assert(MODE == 0 || CK == 2, "partial sums with the K split");
    constexpr int R = ZS64_R, TX = 16, HX = TX + 2, TY = CK == 1 ? 2 * R : R, HY = TY + 2;
    constexpr int CIN = 32 * CK, PAD = CK == 1 ? LDS_PAD32 : LDS_PAD64, PARTS = CIN / 8;
    constexpr int SITES = HX * HY, CHUNKS = SITES * PARTS, PT = (CHUNKS + 511) / 512;
    __shared__ __attribute__((aligned(16))) bf16 zsh[2 * SITES * PAD];  // 62,208 / 57,600 bytes
    __shared__ f32x4 red[CK == 2 ? 2 * 2 * 4 * R * 64 : 1];              // 65,536 bytes at CK 2
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int jb = wave & 3, g = wave >> 2;
    const int kc = CK == 2 ? g : 0, rg = CK == 1 ? g : 0;  // K half, row group
    int r = blockIdx.x;
    const int tx = r % ntx;
    r /= ntx;
    const int ty = r % nty;
    r /= nty;
    const int nzc = D / zc;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = tx * TX, y0 = ty * TY, z0 = zci * zc;
    const int xl = lane & 15, q = lane >> 4;
    bf16x8 w[27];
#pragma unroll
    for (int tap = 0; tap < 27; tap++)
        w[tap] = *(const bf16x8 *)(Wt + ((size_t)tap * wcout + co0 + 16 * jb + xl) * wld + woff + 32 * kc + 8 * q);
    const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
    const f32x4 bias0 = PIN ? zero4 : bias4(bias, co0 + 16 * jb + 4 * q);
    const f32x4 init = kc == 0 ? bias0 : zero4;
    unsigned fo[PT];
#pragma unroll
    for (int i = 0; i < PT; i++) {
        const int c = min(tid + 512 * i, CHUNKS - 1);
        const int s = c / PARTS, part = c % PARTS;
        const int hy = s / HX, hx = s - hy * HX;
        fo[i] = (unsigned)((wrapm(y0 + hy - 1, H) * Wd + wrapm(x0 + hx - 1, Wd)) * ldx + xoff + 8 * part);
    }
    bf16x8 pf[PT];
    auto load_plane = [&](int zz) {
        const bf16 *pl = X + ((size_t)b * D + wrapm(zz, D)) * H * (size_t)Wd * ldx;
#pragma unroll
        for (int i = 0; i < PT; i++)
            pf[i] = *(const bf16x8 *)(pl + fo[i]);
    };
    auto store_plane = [&](int buf) {
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const int c = min(tid + 512 * i, CHUNKS - 1);
            *(bf16x8 *)(zsh + ((size_t)buf * SITES + c / PARTS) * PAD + 8 * (c % PARTS)) = pf[i];
        }
    };
    auto out_at = [&](int zz, int row) {
        return ((((size_t)b * D + zz) * H + y0 + row) * Wd + x0 + xl) * ldy + yoff + 16 * jb + 4 * q;
    };
    // output plane zz (accumulators S, plane index k): CK = 1 stores its
    // rows; CK = 2 writes its half of K to LDS buffer k & 1, and after the
    // next plane's barrier combine() adds both halves of half of the rows
    auto finish = [&](const f32x4(&S)[R], int zz, int k) {
        if constexpr (CK == 1) {
#pragma unroll
            for (int rr = 0; rr < R; rr++)
                *(bf16x4 *)(Y + out_at(zz, R * rg + rr)) = leaky4(S[rr]);
        } else {
            f32x4 *const mine = red + (((k & 1) * 8 + g * 4 + jb) * R) * 64 + lane;
#pragma unroll
            for (int rr = 0; rr < R; rr++)
                mine[rr * 64] = S[rr];
        }
    };
    // (PIN: the partial sums of this wave's combine rows of output zz, a
    // plane ahead)
    f32x4 pin[R / 2];
    auto pin_load = [&](int zz) {
        if constexpr (PIN) {
#pragma unroll
            for (int i = 0; i < R / 2; i++) {
                const int rr = (R / 2) * g + i;
                pin[i] = *(const f32x4 *)(P + ((((size_t)b * D + zz) * H + y0 + rr) * Wd + x0 + xl) * pld +
                                          co0 + 16 * jb + 4 * q);
            }
        }
    };
    // (POOL: the running pooling sums of this wave's row pair)
    float pp[4];
    auto combine = [&](int zz, int k) {
        if constexpr (CK == 2) {
            const f32x4 *const h0 = red + (((k & 1) * 8 + jb) * R) * 64 + lane;
            const f32x4 *const h1 = red + (((k & 1) * 8 + 4 + jb) * R) * 64 + lane;
            bf16x4 o4[R / 2];
#pragma unroll
            for (int i = 0; i < R / 2; i++) {
                const int rr = (R / 2) * g + i;
                if (PIN) {
                    *(bf16x4 *)(Y + out_at(zz, rr)) = leaky4((pin[i] + h0[rr * 64]) + h1[rr * 64]);
                } else if (POUT) {
                    *(f32x4 *)(P + ((((size_t)b * D + zz) * H + y0 + rr) * Wd + x0 + xl) * pld + co0 + 16 * jb +
                               4 * q) = h0[rr * 64] + h1[rr * 64];
                } else {
                    o4[i] = leaky4(h0[rr * 64] + h1[rr * 64]);
                    *(bf16x4 *)(Y + out_at(zz, rr)) = o4[i];
                }
            }
            if constexpr (POOL) {
                static_assert(R / 2 == 2, "one row pair per wave");
                bf16x4 nb[2];  // the x + 1 neighbour's values (lane xl ^ 1)
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    const uint2 u = __builtin_bit_cast(uint2, o4[i]);
                    uint2 v;
                    v.x = __builtin_amdgcn_mov_dpp(u.x, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
                    v.y = __builtin_amdgcn_mov_dpp(u.y, 0xB1, 0xF, 0xF, false);
                    nb[i] = __builtin_bit_cast(bf16x4, v);
                }
                const bool odd = zz & 1;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    float t = odd ? pp[e] : 0.f;
                    t += (float)o4[0][e];
                    t += (float)nb[0][e];
                    t += (float)o4[1][e];
                    t += (float)nb[1][e];
                    pp[e] = t;
                }
                if (odd && !(xl & 1)) {
                    bf16x4 r4;
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        r4[e] = (bf16)(pp[e] * 0.125f);
                    const size_t pv = (((size_t)b * (D >> 1) + (zz >> 1)) * (H >> 1) + ((y0 + 2 * g) >> 1)) *
                                          (Wd >> 1) + ((x0 + xl) >> 1);
                    *(bf16x4 *)(Q + pv * 64 + 16 * jb + 4 * q) = r4;
                }
            }
        }
    };
    const int lofs = ((R * rg) * HX + xl) * PAD + 32 * kc + 8 * q;
    auto plane = [&](auto MSK, f32x4(&Sp)[R], f32x4(&S0)[R], f32x4(&Sm)[R], int k) {
        constexpr int M = decltype(MSK)::value, NR = 3 * (R + 2);
        const bf16 *pb = zsh + (size_t)(k & 1) * SITES * PAD + lofs;
        auto rd = [&](int idx) {
            const int dx = idx / (R + 2), rr = idx % (R + 2);
            return *(const bf16x8 *)(pb + (rr * HX + dx) * PAD);
        };
        bf16x8 a[3];
        a[0] = rd(0);
        a[1] = rd(1);
#pragma unroll
        for (int idx = 0; idx < NR; idx++) {
            if (idx + 2 < NR)
                a[(idx + 2) % 3] = rd(idx + 2);
            __builtin_amdgcn_sched_barrier(0);
            const int dx = idx / (R + 2), rr = idx % (R + 2);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const int ro = rr - dy;
                if (ro >= 0 && ro < R) {
                    const bf16x8 av = a[idx % 3];
                    if (M & 4)
                        Sp[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[dy * 3 + dx], av, Sp[ro], 0, 0, 0);
                    if (M & 2)
                        S0[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[9 + dy * 3 + dx], av, S0[ro], 0, 0, 0);
                    if (M & 1)
                        Sm[ro] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[18 + dy * 3 + dx], av, Sm[ro], 0, 0,
                                                                         0);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // (CB: combine output p - 2, finished by the previous plane)
    auto step = [&](auto MSK, auto STG, auto LDG, auto CB, f32x4(&Sp)[R], f32x4(&S0)[R], f32x4(&Sm)[R],
                    int k) {
        const int p = z0 - 1 + k;
        __syncthreads();  // plane p is staged; plane p - 1's buffer and the older sums are no longer read
        if (decltype(CB)::value)
            combine(p - 2, k - 1);
        if (decltype(STG)::value)
            store_plane((k + 1) & 1);
        if (decltype(MSK)::value & 1)
            pin_load(p - 1);  // for the next combine (PIN)
        if (decltype(LDG)::value)
            load_plane(p + 2);
        plane(MSK, Sp, S0, Sm, k);
        if (decltype(MSK)::value & 1)
            finish(Sm, p - 1, k);
#pragma unroll
        for (int rr = 0; rr < R; rr++)
            Sm[rr] = init;
    };
    f32x4 A0[R], A1[R], A2[R];
#pragma unroll
    for (int rr = 0; rr < R; rr++)
        A0[rr] = A1[rr] = A2[rr] = init;
#pragma unroll
    for (int i = 0; i < 27; i++)
        settle(w[i]);
    settle(init);
    load_plane(z0 - 1);
    store_plane(0);
    load_plane(z0);
    constexpr int SPLIT = CK == 2;
    const int T = (zc + 2) / 3;
    step(IC<4>{}, IC<1>{}, IC<1>{}, IC<0>{}, A0, A2, A1, 0);
    step(IC<6>{}, IC<1>{}, IC<1>{}, IC<0>{}, A1, A0, A2, 1);
    step(IC<7>{}, IC<1>{}, IC<1>{}, IC<0>{}, A2, A1, A0, 2);
    for (int t = 1; t < T - 1; t++) {
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<SPLIT>{}, A0, A2, A1, 3 * t);
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<SPLIT>{}, A1, A0, A2, 3 * t + 1);
        step(IC<7>{}, IC<1>{}, IC<1>{}, IC<SPLIT>{}, A2, A1, A0, 3 * t + 2);
    }
    step(IC<7>{}, IC<1>{}, IC<1>{}, IC<SPLIT>{}, A0, A2, A1, 3 * (T - 1));
    step(IC<3>{}, IC<1>{}, IC<0>{}, IC<SPLIT>{}, A1, A0, A2, 3 * (T - 1) + 1);
    step(IC<1>{}, IC<0>{}, IC<0>{}, IC<SPLIT>{}, A2, A1, A0, 3 * (T - 1) + 2);
    if (SPLIT) {
        __syncthreads();
        combine(z0 + zc - 1, 3 * (T - 1) + 2);
    }
}

// ----------------------------------------------------------------------------
// First layer (3 input channels) without padding each tap to 32 channels.
// K runs over the 9 (dz, dy) groups x 16 slots; slot s = 4 dx' + c (dx' =
// dx + 1 in 0..2, c in 0..2) holds channel c of tap dx, slots with dx' = 3
// or c = 3 are zero: K = 144 in 5 chunks of 32 (the last half empty), 5 MFMA
// chunks per site row instead of 27.  A workgroup owns 64 x 8 sites and
// marches a chunk of z planes (wave w: the 16-site column 16 w .. +15, all 8
// rows); each input plane (tile + one-site halo, 10 x 66 sites) is staged
// once in LDS as BF16, four channels per site (s, l, c, 0: 8 bytes, the same
// rounding as the padded input), for the three output planes that read it.
// A lane's 8 K values of a chunk are then two adjacent sites (slots 0..7:
// dx' = 0, 1) or one site and four zeros (slots 8..15: dx' = 2): one or two
// 8-byte LDS reads.  Weights (5 chunks x Cout x 32, k_prep_w0) stay in
// registers.  Output bf16(leaky(acc + bias)).
// ----------------------------------------------------------------------------
constexpr int C0_TX = 64, C0_R = 8, C0_HX = C0_TX + 2, C0_HY = C0_R + 2, C0_CH = 5;
constexpr int C0_SITES = C0_HY * C0_HX;                   // 660 per plane
constexpr int C0_PT = (C0_SITES + 255) / 256;             // staged sites per thread
constexpr int C0_ZC = 16;                                 // planes per workgroup

// The network's input: packed FP32 [V][3] (InPacked), or a stepping
// context's three field arrays of type T (InFields; the value is rounded to
// FP32 first, as the packed copy would be).
struct InPacked {
    const float *x;
    __device__ float get(size_t v, int ch) const { return x[v * 3 + ch]; }
};
template <typename T>
struct InFields {
    const T *s, *l, *c;
    __device__ float get(size_t v, int ch) const
    {
        return (float)(ch == 0 ? s[v] : (ch == 1 ? l[v] : c[v]));
    }
};

// FULL (the tiles cover the grid, Cout = 16 NJ): every LDS and output store
// is issued unconditionally, so that the compiler can count them in its
// waits (vmcnt counts stores too: a store it cannot count makes the next
// plane's wait for the staging loads a wait for every store).
// (two waves per SIMD for NJ = 2: with the loads fenced a plane ahead the
// compiler would otherwise take 260 registers and one wave; NJ = 4 needs them)
template <int NJ, class Src, bool FULL = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NJ == 2 ? 2 : 1, NJ == 2 ? 2 : 1))) void k_conv0(
    const Src src, const bf16 *__restrict__ W0,
                                               const float *__restrict__ bias, int Cout,
                                               bf16 *__restrict__ Y, int ldy, int B, int D, int H,
                                               int Wd)
{
    __shared__ __attribute__((aligned(16))) bf16 sx[3 * C0_SITES * 4];  // 15,840 bytes
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nbx = (Wd + C0_TX - 1) / C0_TX, nby = (H + C0_R - 1) / C0_R;
    const int nzc = (D + C0_ZC - 1) / C0_ZC;
    int r = blockIdx.x;
    const int bx = r % nbx;
    r /= nbx;
    const int by = r % nby;
    r /= nby;
    const int zci = r % nzc, b = r / nzc;
    const int x0 = bx * C0_TX, y0 = by * C0_R, z0 = zci * C0_ZC, zend = min(z0 + C0_ZC, D);
    // staging: site i = tid + 256 k (row i / C0_HX, x i % C0_HX); threads
    // past the plane re-read its last site and do not store
    unsigned so[C0_PT];
#pragma unroll
    for (int k = 0; k < C0_PT; k++) {
        const int i = min(tid + 256 * k, C0_SITES - 1);
        const int hy = i / C0_HX, hx = i - hy * C0_HX;
        so[k] = (unsigned)(wrapm(y0 + hy - 1, H) * Wd + wrapm(x0 + hx - 1, Wd));
    }
    const size_t plane = (size_t)H * Wd;
    float pv[C0_PT][3];
    auto load = [&](int zz) {
        const size_t base = ((size_t)b * D + wrapm(zz, D)) * plane;
#pragma unroll
        for (int k = 0; k < C0_PT; k++)
#pragma unroll
            for (int ch = 0; ch < 3; ch++)
                pv[k][ch] = src.get(base + so[k], ch);
    };
    auto store = [&](int slot) {
#pragma unroll
        for (int k = 0; k < C0_PT; k++) {
            // (FULL: the threads past the plane rewrite its last site with
            // its own value)
            const int i = FULL ? min(tid + 256 * k, C0_SITES - 1) : tid + 256 * k;
            if (FULL || i < C0_SITES) {
                bf16x4 v;
                v[0] = (bf16)pv[k][0];
                v[1] = (bf16)pv[k][1];
                v[2] = (bf16)pv[k][2];
                v[3] = (bf16)0.0f;
                *(bf16x4 *)(sx + ((size_t)slot * C0_SITES + i) * 4) = v;
            }
        }
    };
    const int m = lane & 15, q = lane >> 4;
    bf16x8 w[C0_CH][NJ];
#pragma unroll
    for (int c = 0; c < C0_CH; c++)
#pragma unroll
        for (int j = 0; j < NJ; j++)
            w[c][j] = *(const bf16x8 *)(W0 + ((size_t)c * Cout + 16 * j + m) * 32 + 8 * q);
    f32x4 bj[NJ];
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int c0 = 16 * j + 4 * q;
        bj[j] = c0 < Cout ? bias4(bias, c0) : f32x4{0.f, 0.f, 0.f, 0.f};
    }
    // plane P in slot (P - z0 + 1) % 3
    load(z0 - 1);
    store(0);
    load(z0);
    store(1);
    load(z0 + 1);
    const int sxl = wave * 16 + m;  // staged x of the site's dx = -1 neighbour
    const int xx = x0 + sxl;
    for (int z = z0; z < zend; z++) {
        __syncthreads();  // every read of the previous plane is done
        store((z - z0 + 2) % 3);
        __syncthreads();
        load(z + 2);  // the next plane's (past the chunk: read and not used)
        // (the fence keeps the loads here, a whole plane ahead of their use;
        // left alone the compiler issues them after most of the MFMAs)
        __builtin_amdgcn_sched_barrier(0);
        f32x4 acc[C0_R][NJ];
#pragma unroll
        for (int row = 0; row < C0_R; row++)
#pragma unroll
            for (int j = 0; j < NJ; j++)
                acc[row][j] = bj[j];
#pragma unroll
        for (int c = 0; c < C0_CH; c++) {
            const int g = 2 * c + (q >> 1);  // (dz, dy) group of this lane's 8 K values
            const int gg = min(g, 8);
            const int dz = gg / 3, dy = gg - 3 * (gg / 3);
            const bf16 *pp = sx + (((size_t)((z - z0 + dz) % 3) * C0_HY + dy) * C0_HX + sxl + 2 * (q & 1)) * 4;
#pragma unroll
            for (int row = 0; row < C0_R; row++) {
                const bf16 *pr = pp + row * C0_HX * 4;
                // (odd q reads its one site twice: the site after it may lie
                // past the staged planes)
                bf16x4 lo = *(const bf16x4 *)pr, hi = *(const bf16x4 *)(pr + 4 * (1 - (q & 1)));
                if (q & 1)
                    hi = bf16x4{(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
                if (g >= 9)
                    lo = hi = bf16x4{(bf16)0.0f, (bf16)0.0f, (bf16)0.0f, (bf16)0.0f};
                const bf16x8 a = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    acc[row][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[c][j], a, acc[row][j], 0, 0, 0);
            }
        }
        // lane holds channels 16j+4q .. +3 of site (x0+16w+m, y0+row) in acc[row][j]
        if (FULL || xx < Wd) {
#pragma unroll
            for (int row = 0; row < C0_R; row++) {
                const int y = y0 + row;
                if (!FULL && y >= H)
                    break;
                const size_t vox = (((size_t)b * D + z) * H + y) * Wd + xx;
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    const int c0 = 16 * j + 4 * q;
                    if (FULL || c0 < Cout)
                        *(bf16x4 *)(Y + vox * ldy + c0) = leaky4(acc[row][j]);
                }
            }
        }
    }
}

// k_conv0's weights: W0[c][co][kk] = w[tap][co][ci] for K index 32c + kk =
// 16 g + s with g = (dz+1)*3 + (dy+1) < 9 and s = 4 (dx+1) + ci, dx+1 < 3,
// ci < 3; else 0.  w is the first layer's flat [27][Cout][3].
__global__ __launch_bounds__(256) void k_prep_w0(const float *__restrict__ w, int Cout,
                                                 bf16 *__restrict__ W0)
{
    const int n = C0_CH * Cout * 32;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const int kk = i % 32, co = (i / 32) % Cout, c = i / (32 * Cout);
        const int k = 32 * c + kk, g = k / 16, s = k % 16;
        float v = 0.f;
        if (g < 9 && s / 4 < 3 && s % 4 < 3) {
            const int tap = g * 3 + s / 4, ci = s % 4;
            v = w[((size_t)tap * Cout + co) * 3 + ci];
        }
        W0[i] = (bf16)v;
    }
}

// k_conv_poly's weights for the first cup (upsampled) input channels of a
// layer: Wp[par][grp][ci / 32][co][ci % 32], par = 4 pz + 2 py + px, grp =
// 4 oz + 2 oy + ox, is the sum of the FP32 weights w[tap][co][ci] over the
// taps (dz, dy, dx) with floor((p + d) / 2) = p - 1 + o in every dimension,
// rounded to BF16 once.  An MFMA fragment (16 output channels x 32 input
// channels) is one contiguous kilobyte: its load touches eight whole cache
// lines instead of halves of sixteen.
__global__ __launch_bounds__(256) void k_prep_wpoly(const float *__restrict__ w, int Cout,
                                                    int CinLog, int cup, bf16 *__restrict__ Wp)
{
    const size_t n = (size_t)64 * Cout * cup;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int ci = (int)((i / ((size_t)32 * Cout)) % (cup / 32)) * 32 + (int)(i % 32);
        const int co = (int)((i / 32) % Cout);
        const int grp = (int)((i / ((size_t)cup * Cout)) % 8);
        const int par = (int)(i / ((size_t)8 * cup * Cout));
        const int p[3] = {(par >> 2) & 1, (par >> 1) & 1, par & 1};
        const int o[3] = {(grp >> 2) & 1, (grp >> 1) & 1, grp & 1};
        float s = 0.f;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    const int d[3] = {dz, dy, dx};
                    bool hit = true;
                    for (int a = 0; a < 3; a++)  // floor((p + d) / 2) - (p - 1) == o
                        hit = hit && ((p[a] + d[a] + 2) / 2 - 1 - p[a] + 1 == o[a]);
                    if (hit) {
                        const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                        s += w[((size_t)tap * Cout + co) * CinLog + ci];
                    }
                }
        Wp[i] = (bf16)s;
    }
}

// ----------------------------------------------------------------------------
// Weight gradient partials: part[split][tap][co][ci] = sum over the split's
// sites of G[v][co] * X[v+s_tap][ci].  A = G^T (16 co x 32 sites), B = the
// shifted input (32 sites x 16 ci), gathered per lane.  Workgroup: one tap,
// 64 output channels (16 per wave), 32 input channels (two B tiles).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad(const bf16 *__restrict__ G, int ldg, int Cout,
                                               const bf16 *__restrict__ X, int ldx, int xoff,
                                               int CinStore, float *__restrict__ part, int nsplit,
                                               size_t vps, int B, int D, int H, int Wd)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tap = blockIdx.x / nsplit, split = blockIdx.x % nsplit;
    const int co0 = blockIdx.y * 64 + wave * 16;
    const int ci0 = blockIdx.z * 32;
    if (co0 >= Cout)
        return;  // whole wave; no workgroup barrier follows
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const size_t V = (size_t)D * H * Wd, total = V * B;
    const size_t vstart = (size_t)split * vps;
    const size_t vend = std::min(total, vstart + vps);
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
    const int co = co0 + (lane & 15);
    const int ci = ci0 + (lane & 15);
    for (size_t vb = vstart; vb < vend; vb += 32) {
        bf16x8 a, b0, b1;
        size_t v0 = vb + (lane >> 4) * 8;
        Coord c = coord_of(v0 < total ? v0 : 0, V, H, Wd);
        for (int j = 0; j < 8; j++) {
            const size_t v = v0 + j;
            if (v < vend) {
                a[j] = G[v * ldg + co];
                const size_t nv = (size_t)c.b * V +
                                  ((size_t)wrapi(c.z + dz, D) * H + wrapi(c.y + dy, H)) * Wd +
                                  wrapi(c.x + dx, Wd);
                b0[j] = X[nv * ldx + xoff + ci];
                b1[j] = X[nv * ldx + xoff + ci + 16];
            } else {
                a[j] = (bf16)0.0f;
                b0[j] = (bf16)0.0f;
                b1[j] = (bf16)0.0f;
            }
            // next site in x-fastest order
            if (++c.x == Wd) {
                c.x = 0;
                if (++c.y == H) {
                    c.y = 0;
                    if (++c.z == D) {
                        c.z = 0;
                        ++c.b;
                    }
                }
            }
        }
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b1, acc1, 0, 0, 0);
    }
    for (int r = 0; r < 4; r++) {
        const int o = co0 + (lane >> 4) * 4 + r;
        if (o >= Cout)
            continue;
        float *p = part + (((size_t)split * 27 + tap) * Cout + o) * CinStore;
        const int i0 = ci0 + (lane & 15);
        if (i0 < CinStore)
            p[i0] = acc0[r];
        if (i0 + 16 < CinStore)
            p[i0 + 16] = acc1[r];
    }
}

// out[i] = sum_s part[s*n + i], fixed order.
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ part, int nsplit,
                                                    size_t n, float *__restrict__ out)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float s = 0.f;
        for (int k = 0; k < nsplit; k++)
            s += part[(size_t)k * n + i];
        out[i] = s;
    }
}

// Per-block column sums of G[v][c] (bias gradients): part[block][c].
__global__ __launch_bounds__(256) void k_colsum(const bf16 *__restrict__ G, int ld, int C,
                                                size_t total, size_t chunk,
                                                float *__restrict__ part)
{
    const size_t v0 = (size_t)blockIdx.x * chunk, v1 = std::min(total, v0 + chunk);
    for (int c = threadIdx.x; c < C; c += 256) {
        float s = 0.f;
        for (size_t v = v0; v < v1; v++)
            s += (float)G[v * ld + c];
        part[(size_t)blockIdx.x * C + c] = s;
    }
}

// Gb[v][c] = bf16(g[v][c] * leaky'(y[v][c])); y > 0 exactly when the
// pre-activation is, since the slope is positive.
__global__ __launch_bounds__(256) void k_act_bwd(const float *__restrict__ g, int ldg, int goff,
                                                 const bf16 *__restrict__ y, int ldy, int yoff,
                                                 int C, size_t total, bf16 *__restrict__ Gb)
{
    const size_t n = total * C;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t v = i / C;
        const int c = (int)(i - v * C);
        const float d = (float)y[v * ldy + yoff + c] > 0.f ? 1.f : LEAK;
        Gb[i] = (bf16)(g[v * ldg + goff + c] * d);
    }
}

// Forward glue kernels.  One thread moves 8 channels (16 bytes); a block
// covers part of one grid row, so the only divisions are per block and per
// channel octet.  Channel counts, strides and offsets are multiples of 8.
__device__ __forceinline__ void row_split(int nch, int &row, int &t)
{
    row = blockIdx.x / nch;
    t = (blockIdx.x - row * nch) * 256 + threadIdx.x;
}

// 2x2x2 average pooling of channels [soff, soff+C) of S (row stride lds) on
// the grid [B][D][H][W] into P [B][D/2][H/2][W/2][C]; the 8 fine sites are
// summed in (dz, dy, dx) order in FP32 and scaled by 1/8.
__global__ __launch_bounds__(256) void k_pool_fwd(const bf16 *__restrict__ S, int lds, int soff,
                                                  int C, bf16 *__restrict__ P, int B, int D,
                                                  int H, int W)
{
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2, oct = C >> 3;
    const int per_row = W2 * oct;
    int row, t;
    row_split((per_row + 255) / 256, row, t);
    if (t >= per_row)
        return;
    const int x2 = t / oct, o = t - x2 * oct;
    const int y2 = row % H2, bz = row / H2, z2 = bz % D2, b = bz / D2;
    float s[8];
    for (int e = 0; e < 8; e++)
        s[e] = 0.f;
    for (int k = 0; k < 8; k++) {
        const size_t v =
            ((size_t)(b * D + 2 * z2 + (k >> 2)) * H + 2 * y2 + ((k >> 1) & 1)) * W + 2 * x2 + (k & 1);
        const bf16x8 a = *(const bf16x8 *)(S + v * lds + soff + 8 * o);
        for (int e = 0; e < 8; e++)
            s[e] += (float)a[e];
    }
    bf16x8 r;
    for (int e = 0; e < 8; e++)
        r[e] = (bf16)(s[e] * 0.125f);
    *(bf16x8 *)(P + ((size_t)row * W2 + x2) * C + 8 * o) = r;
}

// Nearest x2 upsampling of S [B][D/2][H/2][W/2][C] into channels
// [uoff, uoff+C) of U (row stride ldu) on the grid [B][D][H][W].
__global__ __launch_bounds__(256) void k_up_fwd(const bf16 *__restrict__ S, int C,
                                                bf16 *__restrict__ U, int ldu, int uoff, int B,
                                                int D, int H, int W)
{
    const int oct = C >> 3, per_row = W * oct;
    int row, t;
    row_split((per_row + 255) / 256, row, t);
    if (t >= per_row)
        return;
    const int x = t / oct, o = t - x * oct;
    const int y = row % H, bz = row / H, z = bz % D, b = bz / D;
    const size_t src = ((size_t)(b * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2;
    *(bf16x8 *)(U + ((size_t)row * W + x) * ldu + uoff + 8 * o) =
        *(const bf16x8 *)(S + src * C + 8 * o);
}

// Head straight into a stepping context: the residual input is the
// context's own fields, and out = x + head replaces them in place (each site
// is read and written by its own lane only).  head_mfma, as k_head_fwd.
template <typename T, int NK>
__global__ __launch_bounds__(256) void k_head_ctx(const bf16 *__restrict__ F,
                                                  const bf16 *__restrict__ Whd,
                                                  const float *__restrict__ hb, size_t total,
                                                  T *__restrict__ s, T *__restrict__ l,
                                                  T *__restrict__ c)
{
    const int lane = threadIdx.x & 63, xl = lane & 15, q = lane >> 4;
    const size_t v0 = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (v0 >= total)
        return;  // whole wave
    const size_t v = v0 + xl, vr = v < total ? v : total - 1;
    bf16x8 f[NK];
#pragma unroll
    for (int k = 0; k < NK; k++)
        f[k] = *(const bf16x8 *)(F + vr * (32 * NK) + 32 * k + 8 * q);
    const f32x4 o = head_mfma<NK>(Whd, hb, f, lane);
    if (q == 0 && v < total) {
        s[v] = (T)((float)s[v] + o[0]);
        l[v] = (T)((float)l[v] + o[1]);
        c[v] = (T)((float)c[v] + o[2]);
    }
}

// k_mass_delta with the target read from the device (the pre-jump sum).
__global__ __launch_bounds__(256) void k_mass_delta_dev(const double *__restrict__ p, int nb,
                                                        const double *__restrict__ target,
                                                        double n, double *__restrict__ out)
{
    k_mass_delta_body(p, nb, target[0], n, out);
}

// Network input: x [total][3] FP32 -> X0 [total][32] BF16, channels 3..31 zero.
__global__ __launch_bounds__(256) void k_pack_in(const float *__restrict__ x, size_t total,
                                                 bf16 *__restrict__ X0)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total * 4)
        return;
    const size_t v = i >> 2;
    bf16x8 r;
    for (int e = 0; e < 8; e++)
        r[e] = (bf16)0.0f;
    if ((i & 3) == 0) {
        r[0] = (bf16)x[v * 3];
        r[1] = (bf16)x[v * 3 + 1];
        r[2] = (bf16)x[v * 3 + 2];
    }
    *(bf16x8 *)(X0 + i * 8) = r;
}

// Head: out = x + hb + hw F (1x1x1, 32 NK -> 3 channels) by head_mfma, a
// wave per 16 sites.
template <int NK>
__global__ __launch_bounds__(256) void k_head_fwd(const bf16 *__restrict__ F,
                                                  const bf16 *__restrict__ Whd,
                                                  const float *__restrict__ hb,
                                                  const float *__restrict__ x, size_t total,
                                                  float *__restrict__ out)
{
    const int lane = threadIdx.x & 63, xl = lane & 15, q = lane >> 4;
    const size_t v0 = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (v0 >= total)
        return;  // whole wave
    const size_t v = v0 + xl, vr = v < total ? v : total - 1;
    bf16x8 f[NK];
#pragma unroll
    for (int k = 0; k < NK; k++)
        f[k] = *(const bf16x8 *)(F + vr * (32 * NK) + 32 * k + 8 * q);
    const f32x4 o = head_mfma<NK>(Whd, hb, f, lane);
    if (q == 0 && v < total) {
        out[v * 3 + 0] = x[v * 3 + 0] + o[0];
        out[v * 3 + 1] = x[v * 3 + 1] + o[1];
        out[v * 3 + 2] = x[v * 3 + 2] + o[2];
    }
}

// gS[child][c] += gP[parent][c] / 8
__global__ __launch_bounds__(256) void k_pool_bwd(const float *__restrict__ gP, int C,
                                                  float *__restrict__ gS, int lds, int soff,
                                                  int B, int D, int H, int W)
{
    const size_t V = (size_t)D * H * W, n = V * B * C;
    const int H2 = H / 2, W2 = W / 2, D2 = D / 2;
    const size_t V2 = (size_t)D2 * H2 * W2;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t v = i / C;
        const int c = (int)(i - v * C);
        const Coord q = coord_of(v, V, H, W);
        const size_t p = (size_t)q.b * V2 + ((size_t)(q.z >> 1) * H2 + (q.y >> 1)) * W2 + (q.x >> 1);
        gS[v * lds + soff + c] += gP[p * C + c] * 0.125f;
    }
}

// gS[parent][c] = sum over the 8 children of gU[child][c]
__global__ __launch_bounds__(256) void k_up_bwd(const float *__restrict__ gU, int ldu, int uoff,
                                                int C, float *__restrict__ gS, int B, int D,
                                                int H, int W)
{
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2;
    const size_t V2 = (size_t)D2 * H2 * W2, n = V2 * B * C;
    const size_t V = (size_t)D * H * W;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t p = i / C;
        const int c = (int)(i - p * C);
        const Coord q = coord_of(p, V2, H2, W2);
        float s = 0.f;
        for (int k = 0; k < 8; k++) {
            const size_t v = (size_t)q.b * V +
                             ((size_t)(2 * q.z + (k >> 2)) * H + 2 * q.y + ((k >> 1) & 1)) * W +
                             2 * q.x + (k & 1);
            s += gU[v * ldu + uoff + c];
        }
        gS[i] = s;
    }
}

// gOut = 2 (out - y) / N; per-block loss partials; gF[v][ci] = sum_c gOut W[c][ci];
// head partials: part[block][c*CB + ci] and part[block][3*CB + c].
__global__ __launch_bounds__(256) void k_head_bwd(const float *__restrict__ out,
                                                  const float *__restrict__ y,
                                                  const bf16 *__restrict__ F, int CB,
                                                  const float *__restrict__ hw, size_t total,
                                                  size_t chunk, float inv_n,
                                                  float *__restrict__ gF,
                                                  float *__restrict__ part,
                                                  float *__restrict__ lossp)
{
    __shared__ float s_g[256][3];
    __shared__ float s_red[256];
    const size_t v0 = (size_t)blockIdx.x * chunk, v1 = std::min(total, v0 + chunk);
    const int np = 3 * CB + 3;
    float acc = 0.f;   // thread t < np accumulates one head partial
    float loss = 0.f;
    for (size_t base = v0; base < v1; base += 256) {
        const size_t v = base + threadIdx.x;
        float g[3] = {0.f, 0.f, 0.f};
        if (v < v1) {
            for (int c = 0; c < 3; c++) {
                const float d = out[v * 3 + c] - y[v * 3 + c];
                loss += d * d;
                g[c] = 2.f * d * inv_n;
            }
            for (int ci = 0; ci < CB; ci++)
                gF[v * CB + ci] = g[0] * hw[ci] + g[1] * hw[CB + ci] + g[2] * hw[2 * CB + ci];
        }
        s_g[threadIdx.x][0] = g[0];
        s_g[threadIdx.x][1] = g[1];
        s_g[threadIdx.x][2] = g[2];
        __syncthreads();
        const int t = threadIdx.x;
        if (t < np) {
            const int cnt = (int)std::min<size_t>(256, v1 - base);
            if (t < 3 * CB) {
                const int c = t / CB, ci = t - c * CB;
                for (int k = 0; k < cnt; k++)
                    acc += s_g[k][c] * (float)F[(base + k) * CB + ci];
            } else {
                const int c = t - 3 * CB;
                for (int k = 0; k < cnt; k++)
                    acc += s_g[k][c];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < np)
        part[(size_t)blockIdx.x * np + threadIdx.x] = acc;
    s_red[threadIdx.x] = loss;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            s_red[threadIdx.x] += s_red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        lossp[blockIdx.x] = s_red[0];
}

// Adam on the flat FP32 parameters.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v,
                                              size_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// BF16 weights of one 3x3x3 layer, [27][Cout][CinPad], and the flipped,
// channel-transposed copy for the input gradient, [27][CinPad][Cout].
__global__ __launch_bounds__(256) void k_prep_w(const float *__restrict__ w, int Cout, int CinLog,
                                                int CinPad, bf16 *__restrict__ wb,
                                                bf16 *__restrict__ wt, bf16 *__restrict__ wc)
{
    const size_t n = (size_t)27 * Cout * CinPad;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const int ci = (int)(i % CinPad);
        const int co = (int)((i / CinPad) % Cout);
        const int tap = (int)(i / ((size_t)CinPad * Cout));
        const float val = ci < CinLog ? w[((size_t)tap * Cout + co) * CinLog + ci] : 0.f;
        wb[i] = (bf16)val;
        wt[((size_t)(26 - tap) * CinPad + ci) * Cout + co] = (bf16)val;
        wc[(((size_t)tap * (CinPad / 32) + ci / 32) * Cout + co) * 32 + ci % 32] = (bf16)val;
    }
}

// Pack a stepping context's current fields into x [V][3] (FP32).
template <typename T>
__global__ __launch_bounds__(256) void k_ctx_to_x(const T *__restrict__ s, const T *__restrict__ l,
                                                  const T *__restrict__ c, size_t n,
                                                  float *__restrict__ x)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        x[3 * q] = (float)s[q];
        x[3 * q + 1] = (float)l[q];
        x[3 * q + 2] = (float)c[q];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_y_to_ctx(const float *__restrict__ y, size_t n,
                                                  T *__restrict__ s, T *__restrict__ l,
                                                  T *__restrict__ c)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        s[q] = (T)y[3 * q];
        l[q] = (T)y[3 * q + 1];
        c[q] = (T)y[3 * q + 2];
    }
}

// c += d[0] in FP64 (the solute projection of an FP64 context)
__global__ __launch_bounds__(256) void k_add_delta(double *__restrict__ c, size_t n,
                                                   const double *__restrict__ d)
{
    const double v = d[0];
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        c[q] += v;
}

struct Layer {
    int cin_log, cin, cout;
    size_t w_off, b_off;   // in the flat FP32 parameters
    size_t wb_off;         // in the BF16 weight arrays
};

unsigned grid_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 65536)); }
// one thread per item, no grid-stride loop
unsigned blocks_for(size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); }
// rows of per_row items each, a whole number of blocks per row (row_split)
unsigned row_blocks(int rows, int per_row) { return (unsigned)rows * (unsigned)((per_row + 255) / 256); }

}  // namespace

struct ppf_unet {
    int base, device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    Layer L[NCONV];
    size_t head_w, head_b;
    int64_t np;
    size_t nwb;
    float *p, *g, *m, *v;   // FP32 parameters, gradient, Adam moments
    bf16 *wb, *wt;          // BF16 weights and their flipped transposes
    bf16 *wc;               // wb in fragment-contiguous order (k_prep_w)
    bf16 *w0;               // first layer's weights for k_conv0, [5][Cout][32]
    bf16 *wpoly;            // k_conv_poly's weights of L6 and L8, [8][8][cup/32][Cout][32] each
    bf16 *whd;              // head weights split into BF16 hi + lo (k_prep_head)
    size_t wp_off6, wp_off8;
    int64_t t;              // Adam steps
    // workspace for one (batch, grid) shape
    int wsB, wsD, wsH, wsW;
    std::vector<void *> ws;
    bf16 *X0, *E1a, *U1, *P1, *E2a, *U2, *P2, *E3a, *E3b, *D2a, *D2b, *D1a, *D1b, *Gb;
    float *xin, *yout, *ytgt;
    float *gD1b, *gD1a, *gU1, *gD2b, *gD2a, *gU2, *gE3b, *gE3a, *gP2, *gE2a, *gP1, *gE1a;
    float *part;
    size_t part_cap;
    double last_ms, last_flop;
    bool timing_pending;    // ev0/ev1 bracket an apply not yet read back
    hipEvent_t evc, evd;    // ordering with a stepping context's stream (ppf_unet_apply)
    double *tmp;            // solute sums of ppf_unet_apply, 2 * 1024 + 8 doubles
    int conv_kind;  // 2-7: row-blocked / z-marching variants (6, 7: polyphase decoder), 1: LDS, 0: direct
    char err[512];
};

static thread_local char g_unet_err[512];  // last create failure, per host thread

static int uerr(ppf_unet *n, int code, const char *fmt, ...)
{
    char *dst = n ? n->err : g_unet_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, 512, fmt, ap);
    va_end(ap);
    return code;
}

// (a failed call's error is also cleared from the thread's last error, so
// later calls do not report it)
#define UHIP(n, call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            (void)hipGetLastError();                                                  \
            return uerr((n), PPF_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
        }                                                                             \
    } while (0)

static void free_ws(ppf_unet *n)
{
    for (void *p : n->ws)
        if (p)
            hipFree(p);
    n->ws.clear();
    n->wsB = n->wsD = n->wsH = n->wsW = 0;
}

template <typename T>
static int ws_alloc(ppf_unet *n, T **ptr, size_t count)
{
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return uerr(n, PPF_ENOMEM, "workspace allocation of %zu bytes failed", count * sizeof(T));
    }
    n->ws.push_back(p);
    *ptr = (T *)p;
    return PPF_OK;
}

static int ensure_ws(ppf_unet *n, int B, int D, int H, int W, bool train)
{
    if (n->wsB == B && n->wsD == D && n->wsH == H && n->wsW == W && (!train || n->gD1b))
        return PPF_OK;
    UHIP(n, hipStreamSynchronize(n->stream));
    free_ws(n);
    n->gD1b = nullptr;
    const int C = n->base;
    const size_t V1 = (size_t)B * D * H * W, V2 = V1 / 8, V3 = V1 / 64;
    int rc = PPF_OK;
#define A(ptr, cnt) \
    if (!rc) rc = ws_alloc(n, &n->ptr, (cnt))
    A(X0, V1 * C);
    A(E1a, V1 * C);
    A(U1, V1 * 3 * C);
    A(P1, V2 * C);
    A(E2a, V2 * 2 * C);
    A(U2, V2 * 6 * C);
    A(P2, V3 * 2 * C);
    A(E3a, V3 * 4 * C);
    A(E3b, V3 * 4 * C);
    A(D2a, V2 * 2 * C);
    A(D2b, V2 * 2 * C);
    A(D1a, V1 * C);
    A(D1b, V1 * C);
    A(xin, V1 * 3);
    A(yout, V1 * 3);
    if (train) {
        A(ytgt, V1 * 3);
        A(Gb, std::max(V1 * C, std::max(V2 * 2 * C, V3 * 4 * C)));
        A(gD1b, V1 * C);
        A(gD1a, V1 * C);
        A(gU1, V1 * 3 * C);
        A(gD2b, V2 * 2 * C);
        A(gD2a, V2 * 2 * C);
        A(gU2, V2 * 6 * C);
        A(gE3b, V3 * 4 * C);
        A(gE3a, V3 * 4 * C);
        A(gP2, V3 * 2 * C);
        A(gE2a, V2 * 2 * C);
        A(gP1, V2 * C);
        A(gE1a, V1 * C);
    }
#undef A
    if (rc) {
        free_ws(n);
        n->gD1b = nullptr;
        return rc;
    }
    n->wsB = B;
    n->wsD = D;
    n->wsH = H;
    n->wsW = W;
    return PPF_OK;
}

static int ensure_part(ppf_unet *n, size_t count)
{
    if (count <= n->part_cap)
        return PPF_OK;
    if (n->part)
        UHIP(n, hipFree(n->part));
    n->part = nullptr;
    n->part_cap = 0;
    UHIP(n, hipMalloc(&n->part, count * sizeof(float)));
    n->part_cap = count;
    return PPF_OK;
}

static int prep_weights(ppf_unet *n)
{
    for (int i = 0; i < NCONV; i++) {
        const Layer &l = n->L[i];
        hipLaunchKernelGGL(k_prep_w, dim3(grid_for((size_t)27 * l.cout * l.cin)), dim3(256), 0,
                           n->stream, n->p + l.w_off, l.cout, l.cin_log, l.cin, n->wb + l.wb_off,
                           n->wt + l.wb_off, n->wc + l.wb_off);
    }
    hipLaunchKernelGGL(k_prep_w0, dim3(grid_for((size_t)C0_CH * n->L[0].cout * 32)), dim3(256), 0,
                       n->stream, n->p + n->L[0].w_off, n->L[0].cout, n->w0);
    hipLaunchKernelGGL(k_prep_head, dim3(grid_for((size_t)n->base / 32 * 1024)), dim3(256), 0,
                       n->stream, n->p + n->head_w, n->base, n->whd);
    const int C = n->base;
    const int lay[2] = {6, 8}, cups[2] = {4 * C, 2 * C};
    const size_t offs[2] = {n->wp_off6, n->wp_off8};
    for (int k = 0; k < 2; k++) {
        const Layer &l = n->L[lay[k]];
        hipLaunchKernelGGL(k_prep_wpoly, dim3(grid_for((size_t)64 * l.cout * cups[k])), dim3(256),
                           0, n->stream, n->p + l.w_off, l.cout, l.cin_log, cups[k],
                           n->wpoly + offs[k]);
    }
    UHIP(n, hipGetLastError());
    return PPF_OK;
}

// Launch one 3x3x3 convolution of layer geometry (cin, cout) on the grid
// [B][D][H][W].  Input channel c is X's element xoff + c of a site (rows of
// ldx); with Xc (inference over upsampled channels) channels c < cup come
// from the coarse tensor Xc and channel c >= cup is X's element
// xoff + c - cup.
// z chunk of the input-plane-stationary kernels: 3k + 1 planes (k >= 1)
// that tile D.  64 planes (4% faster than 16 at 256^3: fewer chunk edges)
// while the launch keeps at least 1024 workgroups (four rounds of one per
// CU), else 16, else any such divisor; 0 if there is none.
static int zs_chunk(int D, size_t tiles_per_plane)
{
    if (D % 64 == 0 && tiles_per_plane * (D / 64) >= 1024)
        return 64;
    if (D % 16 == 0)
        return 16;
    for (int zc = std::min(D, 64); zc >= 4; zc--)
        if (zc % 3 == 1 && D % zc == 0)
            return zc;
    return 0;
}

// The input-plane-stationary level-1 kernel with epilogue epi, where D has a
// chunk of 3k + 1 planes and the in-plane offsets fit 32 bits; the head
// epilogues need full tiles.  false: not launched.
template <class EPI>
static bool launch_zs32(ppf_unet *n, const bf16 *X, int ldx, int xoff, const bf16 *Wt,
                        const float *bias, bf16 *Yb, int ldy, int yoff, int B, int D, int H, int W,
                        const EPI &epi)
{
    constexpr int R = ZS_R;
    const int ntx = (W + ZS_TX - 1) / ZS_TX, nty = (H + 4 * R - 1) / (4 * R);
    const int zs = zs_chunk(D, (size_t)B * ntx * nty);
    const bool full = W % ZS_TX == 0 && H % (4 * R) == 0;
    if (!zs || (size_t)H * W * ldx >= (1u << 31) || (!std::is_same<EPI, EpiStore>::value && !full))
        return false;
    if (std::is_same<EPI, EpiPool>::value && (zs % 2 || D % 2))
        return false;
    const unsigned nblk = (unsigned)((size_t)B * (D / zs) * nty * ntx);
    if (full)
        hipLaunchKernelGGL((k_conv_zs32<R, true, EPI>), dim3(nblk), dim3(512), 0, n->stream, X, ldx,
                           xoff, Wt, bias, Yb, ldy, yoff, B, D, H, W, zs, ntx, nty, epi);
    else if constexpr (std::is_same<EPI, EpiStore>::value)
        hipLaunchKernelGGL((k_conv_zs32<R, false, EPI>), dim3(nblk), dim3(512), 0, n->stream, X, ldx,
                           xoff, Wt, bias, Yb, ldy, yoff, B, D, H, W, zs, ntx, nty, epi);
    return true;
}

static void conv(ppf_unet *n, int mode, const bf16 *X, int ldx, int xoff, int cin,
                 const bf16 *Wt, int coutw, const float *bias, int cout, bf16 *Yb, float *Yf,
                 int ldy, int yoff, int B, int D, int H, int W, const bf16 *Xc = nullptr,
                 int ldc = 0, int cup = 0, const bf16 *Wp = nullptr)
{
    const size_t total = (size_t)B * D * H * W;
    // the row-blocked and polyphase kernels read the forward weights in
    // fragment-contiguous order (mode 0 passes wb's layers)
    const bf16 *Wr = mode == 0 ? n->wc + (Wt - n->wb) : Wt;
    if (n->conv_kind == 6 && Xc && Wp && mode == 0 && cin == 96 &&
        ((ldx == 96 && xoff == 64) || (ldx == 32 && xoff == 0)) &&
        cup == 64 && ldc == 64 && coutw == 32 && cout == 32 && ldy % 4 == 0 && yoff == 0 &&
        D % 2 == 0 && H % 2 == 0 && W % 2 == 0 && (size_t)H * W * ldx < (1u << 31)) {
        // L8 of the base-32 network: the eight-wave z-marching polyphase
        // kernel (k_conv_poly's products; in-plane offsets in 32 bits), z
        // chunks of 16 planes (even)
        const int zc = D < 16 ? D : 16, ntx = (W + ZP_TX - 1) / ZP_TX, nty = (H + ZP_TY - 1) / ZP_TY;
        const unsigned nblk = (unsigned)((size_t)B * ((D + zc - 1) / zc) * nty * ntx);
        if (W % ZP_TX == 0 && H % ZP_TY == 0)
            hipLaunchKernelGGL(k_conv_zpoly8<true>, dim3(nblk), dim3(512), 0, n->stream, X, ldx, xoff, Wt,
                               Wp, bias, Yb, ldy, B, D, H, W, Xc, ldc, zc, ntx, nty);
        else
            hipLaunchKernelGGL(k_conv_zpoly8<false>, dim3(nblk), dim3(512), 0, n->stream, X, ldx, xoff, Wt,
                               Wp, bias, Yb, ldy, B, D, H, W, Xc, ldc, zc, ntx, nty);
        return;
    }
    if ((n->conv_kind == 6 || n->conv_kind == 7) && Xc && Wp && mode == 0 && coutw == cout &&
        coutw % 32 == 0 && ldy % 4 == 0 && yoff == 0 && cup % 32 == 0 && cin % 32 == 0) {
        // decoder layer over an upsampled tensor: polyphase form (8 x 16
        // sites x (32 or 64) channels per wave)
        const bool wide = coutw % 64 == 0;
        const int NJ = wide ? 4 : 2, R = 8;
        const int nnb = coutw / (NJ * 16), nxt2 = 2 * ((W + 31) / 32), nyt = (H + R - 1) / R;
        const size_t waves = (size_t)B * D * nyt * nxt2 * nnb;
        const dim3 g3((unsigned)((waves + 3) / 4));
        if (wide)
            hipLaunchKernelGGL((k_conv_poly<8, 4, 1>), g3, dim3(256), 0, n->stream, X, ldx, xoff, cin, Wr,
                               coutw, Wp, bias, cout, Yb, ldy, B, D, H, W, nnb, nxt2, nyt, Xc, ldc,
                               cup);
        else
            hipLaunchKernelGGL((k_conv_poly<8, 2, 2>), g3, dim3(256), 0, n->stream, X, ldx, xoff, cin, Wr,
                               coutw, Wp, bias, cout, Yb, ldy, B, D, H, W, nnb, nxt2, nyt, Xc, ldc,
                               cup);
        return;
    }
    if ((n->conv_kind == 2 || n->conv_kind >= 5) && mode == 0 && !Xc &&
        (cin == 32 || cin == 64) && coutw == 64 && cout == 64 && ldy % 4 == 0 && yoff % 4 == 0) {
        // Cout 64: the 16 x 16 z-marching kernel (same sums as k_conv_rows);
        // z chunks of 8 planes at most, so that a launch has at least ~4
        // workgroups per CU
        // the input-plane-stationary kernel: CK = 1 (same sums) where the
        // level-1 kernel runs, CK = 2 (K split over two waves) with kind 6
        const int ty = cin == 32 ? 2 * ZS64_R : ZS64_R;
        if ((cin == 32 || n->conv_kind == 6) && W % 16 == 0 && H % ty == 0 &&
            (size_t)H * W * ldx < (1u << 31)) {
            const int sx = W / 16, sy = H / ty;
            const int zs = zs_chunk(D, (size_t)B * sx * sy);
            if (zs) {
                const unsigned nb = (unsigned)((size_t)B * (D / zs) * sy * sx);
                if (cin == 32)
                    hipLaunchKernelGGL((k_conv_zs64<1>), dim3(nb), dim3(512), 0, n->stream, X, ldx, xoff, Wt,
                                       bias, Yb, ldy, yoff, B, D, H, W, zs, sx, sy, 64, 0, 32, 0,
                                       (float *)nullptr, 0);
                else
                    hipLaunchKernelGGL((k_conv_zs64<2>), dim3(nb), dim3(512), 0, n->stream, X, ldx, xoff, Wt,
                                       bias, Yb, ldy, yoff, B, D, H, W, zs, sx, sy, 64, 0, 64, 0,
                                       (float *)nullptr, 0);
                return;
            }
        }
        const int ntx = (W + Z64_T - 1) / Z64_T, nty = (H + Z64_T - 1) / Z64_T;
        const int zc = D < 8 ? D : 8;
        const unsigned nblk = (unsigned)((size_t)B * ((D + zc - 1) / zc) * nty * ntx);
        if (cin == 32)
            hipLaunchKernelGGL(k_conv_z64<1>, dim3(nblk), dim3(256), 0, n->stream, X, ldx, xoff, Wt,
                               bias, Yb, ldy, yoff, B, D, H, W, zc, ntx, nty);
        else
            hipLaunchKernelGGL(k_conv_z64<2>, dim3(nblk), dim3(256), 0, n->stream, X, ldx, xoff, Wt,
                               bias, Yb, ldy, yoff, B, D, H, W, zc, ntx, nty);
        return;
    }
    if ((n->conv_kind == 2 || n->conv_kind >= 5) && mode == 0 && !Xc && cin == 32 &&
        coutw == 32 && cout == 32) {
        // level-1 layers: the z-marching kernels (same sums as k_conv_rows);
        // the input-plane-stationary one where D has a chunk of 3k + 1
        // planes (k >= 1)
        if (launch_zs32(n, X, ldx, xoff, Wt, bias, Yb, ldy, yoff, B, D, H, W, EpiStore{}))
            return;
        const int zc = 16, mtx = (W + ZM_TX - 1) / ZM_TX, mty = (H + ZM_TY - 1) / ZM_TY;
        const unsigned nblk = (unsigned)((size_t)B * ((D + zc - 1) / zc) * mty * mtx);
        hipLaunchKernelGGL(k_conv_z32<8>, dim3(nblk), dim3(512), 0, n->stream, X, ldx, xoff, Wt,
                           bias, Yb, ldy, yoff, B, D, H, W, zc, mtx, mty);
        return;
    }
    // (the row-blocked kernels store 4 channels at once: aligned when the
    // channel count, row stride and offset are multiples of 4)
    if (n->conv_kind >= 2 && coutw % 32 == 0 && cout % 4 == 0 && ldy % 4 == 0 && yoff % 4 == 0) {
        const int kind = n->conv_kind >= 5 ? 2 : n->conv_kind;  // 5-7: kind 2's tiles
        // (R, NJ): 8 rows x 2 N tiles; where the channel count is a multiple
        // of 64, 8 x 4 (kind 2, fastest measured: 18.0 ms at 256^3) or 4 x 4
        // (kind 4); kind 3 keeps 8 x 2 everywhere.  NJ*16 must divide coutw
        // (the kernel has no partial N tile); every layer's padded channel
        // count is a multiple of 32.
        const bool wide = coutw % 64 == 0 && kind != 3;
        const int R = wide && kind == 4 ? 4 : 8, NJ = wide ? 4 : 2;
        const int nnb = (coutw + NJ * 16 - 1) / (NJ * 16), nxt = (W + 15) / 16,
                  nyt = (H + R - 1) / R;
        const size_t waves = (size_t)B * D * nyt * nxt * nnb;
        const dim3 g3((unsigned)((waves + 3) / 4));
#define ROWS(RR, JJ, MM)                                                                      \
    hipLaunchKernelGGL((k_conv_rows<RR, JJ, MM, (RR * JJ > 16 ? 1 : 2), false>), g3, dim3(256),   \
                       0, n->stream, X, ldx, xoff, cin, Wr, coutw, bias, cout, Yb, Yf, ldy, yoff, \
                       B, D, H, W, nnb, nxt, nyt, (const bf16 *)nullptr, 0, 0)
#define ROWS_UP(RR, JJ)                                                                        \
    hipLaunchKernelGGL((k_conv_rows<RR, JJ, 0, (RR * JJ > 16 ? 1 : 2), true>), g3, dim3(256), 0,  \
                       n->stream, X, ldx, xoff, cin, Wr, coutw, bias, cout, Yb, Yf, ldy, yoff, B, \
                       D, H, W, nnb, nxt, nyt, Xc, ldc, cup)
        if (Xc) {  // forward only
            if (R == 8 && NJ == 2)
                ROWS_UP(8, 2);
            else if (R == 4)
                ROWS_UP(4, 4);
            else
                ROWS_UP(8, 4);
            return;
        }
        if (R == 8 && NJ == 2) {
            if (mode == 0)
                ROWS(8, 2, 0);
            else if (mode == 1)
                ROWS(8, 2, 1);
            else
                ROWS(8, 2, 2);
        } else if (R == 4) {
            if (mode == 0)
                ROWS(4, 4, 0);
            else if (mode == 1)
                ROWS(4, 4, 1);
            else
                ROWS(4, 4, 2);
        } else {
            if (mode == 0)
                ROWS(8, 4, 0);
            else if (mode == 1)
                ROWS(8, 4, 1);
            else
                ROWS(8, 4, 2);
        }
#undef ROWS
#undef ROWS_UP
        return;
    }
    if (n->conv_kind >= 1) {
        const size_t bricks = (size_t)B * D * ((H + CL_BY - 1) / CL_BY) * ((W + CL_BX - 1) / CL_BX);
        const dim3 g2((unsigned)bricks, (unsigned)((coutw + CL_CT - 1) / CL_CT));
        if (mode == 0)
            hipLaunchKernelGGL(k_conv_lds<0>, g2, dim3(256), 0, n->stream, X, ldx, xoff, cin, Wt,
                               coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
        else if (mode == 1)
            hipLaunchKernelGGL(k_conv_lds<1>, g2, dim3(256), 0, n->stream, X, ldx, xoff, cin, Wt,
                               coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
        else
            hipLaunchKernelGGL(k_conv_lds<2>, g2, dim3(256), 0, n->stream, X, ldx, xoff, cin, Wt,
                               coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
        return;
    }
    const dim3 grid((unsigned)((total + 63) / 64), (unsigned)((coutw + 63) / 64));
    if (mode == 0)
        hipLaunchKernelGGL((k_conv<4, 0, 27>), grid, dim3(256), 0, n->stream, X, ldx, xoff, cin,
                           Wt, coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
    else if (mode == 1)
        hipLaunchKernelGGL((k_conv<4, 1, 27>), grid, dim3(256), 0, n->stream, X, ldx, xoff, cin,
                           Wt, coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
    else
        hipLaunchKernelGGL((k_conv<4, 2, 27>), grid, dim3(256), 0, n->stream, X, ldx, xoff, cin,
                           Wt, coutw, bias, cout, Yb, Yf, ldy, yoff, B, D, H, W);
}

// A stepping context's fields as the network's input and output (apply).
struct CtxIO {
    void *s, *l, *c;
    bool f64;
};

// need_x0: also build the padded BF16 input X0 (the first layer's weight
// gradient reads it); inference with the row-blocked kernels does not.
// io (row-blocked kernels only): read the input from a context's fields and
// write the output back into them instead of xin / yout.
static int forward(ppf_unet *n, int B, int D, int H, int W, bool need_x0,
                   const CtxIO *io = nullptr)
{
    const int C = n->base;
    const size_t V1 = (size_t)B * D * H * W;
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D3 = D / 4, H3 = H / 4, W3 = W / 4;
    const Layer *L = n->L;
    auto wb = [&](int i) { return n->wb + L[i].wb_off; };
    auto bias = [&](int i) { return n->p + L[i].b_off; };
    // level 1
    if (n->conv_kind >= 2) {
        const unsigned nb0 = (unsigned)((size_t)B * ((D + C0_ZC - 1) / C0_ZC) *
                                        ((H + C0_R - 1) / C0_R) * ((W + C0_TX - 1) / C0_TX));
        const bool full0 = W % C0_TX == 0 && H % C0_R == 0 && D % C0_ZC == 0;
#define CONV0(NJ, SRC)                                                                         \
    do {                                                                                       \
        if (full0)                                                                             \
            hipLaunchKernelGGL((k_conv0<NJ, decltype(SRC), true>), dim3(nb0), dim3(256), 0,     \
                               n->stream, SRC, n->w0, bias(0), C, n->E1a, C, B, D, H, W);      \
        else                                                                                   \
            hipLaunchKernelGGL((k_conv0<NJ, decltype(SRC)>), dim3(nb0), dim3(256), 0, n->stream, \
                               SRC, n->w0, bias(0), C, n->E1a, C, B, D, H, W);                 \
    } while (0)
        const InPacked packed{n->xin};
        if (io && io->f64) {
            const InFields<double> src{(const double *)io->s, (const double *)io->l,
                                       (const double *)io->c};
            if (C == 32)
                CONV0(2, src);
            else
                CONV0(4, src);
        } else if (io) {
            const InFields<float> src{(const float *)io->s, (const float *)io->l,
                                      (const float *)io->c};
            if (C == 32)
                CONV0(2, src);
            else
                CONV0(4, src);
        } else if (C == 32) {
            CONV0(2, packed);
        } else {
            CONV0(4, packed);
        }
#undef CONV0
        if (need_x0)
            hipLaunchKernelGGL(k_pack_in, dim3(blocks_for(V1 * 4)), dim3(256), 0, n->stream,
                               n->xin, V1, n->X0);
    } else {
        hipLaunchKernelGGL(k_pack_in, dim3(blocks_for(V1 * 4)), dim3(256), 0, n->stream, n->xin,
                           V1, n->X0);
        // X0 is [V][32] whatever the base (k_pack_in)
        conv(n, 0, n->X0, 32, 0, L[0].cin, wb(0), L[0].cout, bias(0), L[0].cout, n->E1a, nullptr,
             C, 0, B, D, H, W);
    }
    // inference with the row-blocked kernels reads the upsampled channels
    // straight from the coarse tensor and the skip channels from their own
    // contiguous tensor (in U1's and U2's memory); training materializes the
    // concatenation [upsampled | skip] (the weight gradients read it)
    const bool fuse_up = !need_x0 && (n->conv_kind >= 2 && n->conv_kind != 5);
    const int ld1 = fuse_up ? C : 3 * C, off1 = fuse_up ? 0 : 2 * C;
    const int ld2 = fuse_up ? 2 * C : 6 * C, off2 = fuse_up ? 0 : 4 * C;
    // L1 with the pooling in its epilogue where the level-1 kernel runs (base
    // 32), else L1 and k_pool_fwd
    if (!(C == 32 && (n->conv_kind == 2 || n->conv_kind >= 5) &&
          launch_zs32(n, n->E1a, C, 0, wb(1), bias(1), n->U1, ld1, off1, B, D, H, W,
                      EpiPool{n->P1}))) {
        conv(n, 0, n->E1a, C, 0, C, wb(1), C, bias(1), C, n->U1, nullptr, ld1, off1, B, D, H, W);
        hipLaunchKernelGGL(k_pool_fwd, dim3(row_blocks(B * D2 * H2, W2 * C / 8)), dim3(256), 0,
                           n->stream, n->U1, ld1, off1, C, n->P1, B, D, H, W);
    }
    // level 2
    conv(n, 0, n->P1, C, 0, C, wb(2), 2 * C, bias(2), 2 * C, n->E2a, nullptr, 2 * C, 0, B, D2, H2,
         W2);
    // L3 and the level-2 pooling: one kernel where L3 runs the
    // input-plane-stationary kernel with the K split (kind 6) on z chunks
    // that start at even planes
    {
        const int sy = H2 / ZS64_R, sx = W2 / 16;
        const int zs = (n->conv_kind == 6 && C == 32 && W2 % 16 == 0 && H2 % ZS64_R == 0 &&
                        (size_t)H2 * W2 * 2 * C < (1u << 31))
                           ? zs_chunk(D2, (size_t)B * sx * sy)
                           : 0;
        if (zs && zs % 2 == 0 && ld2 % 4 == 0 && off2 % 4 == 0) {
            const unsigned nb = (unsigned)((size_t)B * (D2 / zs) * sy * sx);
            hipLaunchKernelGGL((k_conv_zs64<2, 3>), dim3(nb), dim3(512), 0, n->stream, n->E2a, 2 * C, 0,
                               wb(3), bias(3), n->U2, ld2, off2, B, D2, H2, W2, zs, sx, sy, 64, 0, 64, 0,
                               (float *)nullptr, 0, n->P2);
        } else {
            conv(n, 0, n->E2a, 2 * C, 0, 2 * C, wb(3), 2 * C, bias(3), 2 * C, n->U2, nullptr, ld2, off2,
                 B, D2, H2, W2);
            hipLaunchKernelGGL(k_pool_fwd, dim3(row_blocks(B * D3 * H3, W3 * 2 * C / 8)), dim3(256), 0,
                               n->stream, n->U2, ld2, off2, 2 * C, n->P2, B, D2, H2, W2);
        }
    }
    // level 3
    // the level-3 layers of the base-32 network (kind 6): the 64-channel
    // input-plane-stationary kernel per half of the 128 output channels;
    // L5 (128 input channels) in two passes over 64 of them each, through
    // FP32 partial sums
    const int sx3 = W3 / 16, sy3 = H3 / ZS64_R;
    const int zs3 = (W3 % 16 == 0 && H3 % ZS64_R == 0) ? zs_chunk(D3, (size_t)B * sx3 * sy3 * 2) : 0;
    if (n->conv_kind == 6 && C == 32 && zs3 && (size_t)H3 * W3 * 4 * C < (1u << 31)) {
        int rc = ensure_part(n, (size_t)B * D3 * H3 * W3 * 4 * C);
        if (rc)
            return rc;
        const unsigned nb = (unsigned)((size_t)B * (D3 / zs3) * sy3 * sx3);
        for (int co0 = 0; co0 < 4 * C; co0 += 64)
            hipLaunchKernelGGL((k_conv_zs64<2, 0>), dim3(nb), dim3(512), 0, n->stream, n->P2, 2 * C, 0, wb(4),
                               bias(4), n->E3a, 4 * C, co0, B, D3, H3, W3, zs3, sx3, sy3, 4 * C, co0, 2 * C, 0,
                               (float *)nullptr, 0);
        for (int co0 = 0; co0 < 4 * C; co0 += 64) {
            hipLaunchKernelGGL((k_conv_zs64<2, 2>), dim3(nb), dim3(512), 0, n->stream, n->E3a, 4 * C, 0, wb(5),
                               bias(5), (bf16 *)nullptr, 4 * C, co0, B, D3, H3, W3, zs3, sx3, sy3, 4 * C, co0,
                               4 * C, 0, n->part, 4 * C);
            hipLaunchKernelGGL((k_conv_zs64<2, 1>), dim3(nb), dim3(512), 0, n->stream, n->E3a, 4 * C, 64,
                               wb(5), bias(5), n->E3b, 4 * C, co0, B, D3, H3, W3, zs3, sx3, sy3, 4 * C, co0,
                               4 * C, 64, n->part, 4 * C);
        }
    } else {
        conv(n, 0, n->P2, 2 * C, 0, 2 * C, wb(4), 4 * C, bias(4), 4 * C, n->E3a, nullptr, 4 * C, 0, B,
             D3, H3, W3);
        conv(n, 0, n->E3a, 4 * C, 0, 4 * C, wb(5), 4 * C, bias(5), 4 * C, n->E3b, nullptr, 4 * C, 0, B,
             D3, H3, W3);
    }
    // decoder
    const bf16 *wp6 = n->wpoly + n->wp_off6, *wp8 = n->wpoly + n->wp_off8;
    // L6 of the base-32 network at inference, kind 6, split in two: the
    // upsampled channels (bias + their polyphase sum, 16 rows per wave) into
    // FP32 partial sums, then the skip channels by the input-plane-stationary
    // kernel from those sums (K split over two waves)
    const int sx6 = W2 / 16, sy6 = H2 / ZS64_R;
    const int zs6 = (W2 % 16 == 0 && H2 % ZS64_R == 0) ? zs_chunk(D2, (size_t)B * sx6 * sy6) : 0;
    if (fuse_up && n->conv_kind == 6 && C == 32 && zs6 && (size_t)H2 * W2 * 2 * C < (1u << 31)) {
        int rc = ensure_part(n, (size_t)B * D2 * H2 * W2 * 2 * C);
        if (rc)
            return rc;
        const int nxt2 = 2 * ((W2 + 31) / 32), nyt = (H2 + 15) / 16;
        if (W2 % 32 == 0 && H2 % 64 == 0) {
            // weights shared by the workgroup's four row tiles through LDS
            const dim3 gu((unsigned)((size_t)B * D2 * nxt2 * (H2 / 64)));
            hipLaunchKernelGGL((k_conv_polyup<8, 8>), gu, dim3(512), 0, n->stream, wp6, bias(6), 2 * C, B,
                               D2, H2, W2, n->E3b, 4 * C, 4 * C, n->part);
        } else {
            const dim3 gu((unsigned)(((size_t)B * D2 * nyt * nxt2 + 3) / 4));
            hipLaunchKernelGGL((k_conv_poly<16, 4, 1, true>), gu, dim3(256), 0, n->stream, n->U2, ld2,
                               off2, 6 * C, n->wc + L[6].wb_off, 2 * C, wp6, bias(6), 2 * C,
                               (bf16 *)nullptr, 2 * C, B, D2, H2, W2, 1, nxt2, nyt, n->E3b, 4 * C, 4 * C,
                               n->part);
        }
        const unsigned nb = (unsigned)((size_t)B * (D2 / zs6) * sy6 * sx6);
        hipLaunchKernelGGL((k_conv_zs64<2, 1>), dim3(nb), dim3(512), 0, n->stream, n->U2, ld2, off2,
                           wb(6), bias(6), n->D2a, 2 * C, 0, B, D2, H2, W2, zs6, sx6, sy6, 2 * C, 0, 6 * C,
                           4 * C, n->part, 2 * C);
    } else if (fuse_up) {
        conv(n, 0, n->U2, ld2, off2, 6 * C, wb(6), 2 * C, bias(6), 2 * C, n->D2a, nullptr, 2 * C, 0,
             B, D2, H2, W2, n->E3b, 4 * C, 4 * C, wp6);
    } else {
        hipLaunchKernelGGL(k_up_fwd, dim3(row_blocks(B * D2 * H2, W2 * 4 * C / 8)), dim3(256), 0,
                           n->stream, n->E3b, 4 * C, n->U2, 6 * C, 0, B, D2, H2, W2);
        conv(n, 0, n->U2, 6 * C, 0, 6 * C, wb(6), 2 * C, bias(6), 2 * C, n->D2a, nullptr, 2 * C, 0,
             B, D2, H2, W2);
    }
    conv(n, 0, n->D2a, 2 * C, 0, 2 * C, wb(7), 2 * C, bias(7), 2 * C, n->D2b, nullptr, 2 * C, 0, B,
         D2, H2, W2);
    if (fuse_up) {
        conv(n, 0, n->U1, ld1, off1, 3 * C, wb(8), C, bias(8), C, n->D1a, nullptr, C, 0, B, D, H, W,
             n->D2b, 2 * C, 2 * C, wp8);
    } else {
        hipLaunchKernelGGL(k_up_fwd, dim3(row_blocks(B * D * H, W * 2 * C / 8)), dim3(256), 0,
                           n->stream, n->D2b, 2 * C, n->U1, 3 * C, 0, B, D, H, W);
        conv(n, 0, n->U1, 3 * C, 0, 3 * C, wb(8), C, bias(8), C, n->D1a, nullptr, C, 0, B, D, H,
             W);
    }
    // the last layer with the head in its epilogue (inference, base 32, the
    // kinds that run the level-1 layers z-marching), else L9 and the head
    const float *hb = n->p + n->head_b;
    if (!need_x0 && C == 32 && (n->conv_kind == 2 || n->conv_kind >= 5)) {
        bool done;
        if (io && io->f64)
            done = launch_zs32(n, n->D1a, C, 0, wb(9), bias(9), nullptr, C, 0, B, D, H, W,
                               HeadCtx<double>{n->whd, hb, (double *)io->s, (double *)io->l,
                                               (double *)io->c});
        else if (io)
            done = launch_zs32(n, n->D1a, C, 0, wb(9), bias(9), nullptr, C, 0, B, D, H, W,
                               HeadCtx<float>{n->whd, hb, (float *)io->s, (float *)io->l, (float *)io->c});
        else
            done = launch_zs32(n, n->D1a, C, 0, wb(9), bias(9), nullptr, C, 0, B, D, H, W,
                               HeadPacked{n->whd, hb, n->xin, n->yout});
        if (done) {
            UHIP(n, hipGetLastError());
            return PPF_OK;
        }
    }
    conv(n, 0, n->D1a, C, 0, C, wb(9), C, bias(9), C, n->D1b, nullptr, C, 0, B, D, H, W);
    const dim3 gh((unsigned)((V1 + 63) / 64));
#define HEADK(K)                                                                                  \
    do {                                                                                          \
        if (io && io->f64)                                                                        \
            hipLaunchKernelGGL((k_head_ctx<double, K>), gh, dim3(256), 0, n->stream, n->D1b, n->whd, \
                               hb, V1, (double *)io->s, (double *)io->l, (double *)io->c);         \
        else if (io)                                                                              \
            hipLaunchKernelGGL((k_head_ctx<float, K>), gh, dim3(256), 0, n->stream, n->D1b, n->whd, \
                               hb, V1, (float *)io->s, (float *)io->l, (float *)io->c);            \
        else                                                                                      \
            hipLaunchKernelGGL((k_head_fwd<K>), gh, dim3(256), 0, n->stream, n->D1b, n->whd, hb,   \
                               n->xin, V1, n->yout);                                              \
    } while (0)
    if (C == 32)
        HEADK(1);
    else
        HEADK(2);
#undef HEADK
    UHIP(n, hipGetLastError());
    return PPF_OK;
}

// Weight and bias gradient of layer i into n->g, from Gb (bf16 [sites][cout])
// and the layer input X (ldx, xoff).
static int layer_wgrad(ppf_unet *n, int i, const bf16 *X, int ldx, int xoff, int B, int D, int H,
                       int W)
{
    const Layer &l = n->L[i];
    const size_t total = (size_t)B * D * H * W;
    const int ncit = l.cin / 32, ncot = (l.cout + 63) / 64;
    // splits: enough workgroups to fill the device, at least 32 sites each
    int nsplit = (int)std::max<size_t>(1, std::min<size_t>((total + 31) / 32,
                                                          4096 / (27 * ncit * ncot) + 1));
    size_t vps = ((total + nsplit - 1) / nsplit + 31) / 32 * 32;
    nsplit = (int)((total + vps - 1) / vps);
    const size_t nw = (size_t)27 * l.cout * l.cin_log;
    int rc = ensure_part(n, (size_t)nsplit * nw);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_wgrad, dim3(27 * nsplit, ncot, ncit), dim3(256), 0, n->stream, n->Gb,
                       l.cout, l.cout, X, ldx, xoff, l.cin_log, n->part, nsplit, vps, B, D, H, W);
    hipLaunchKernelGGL(k_sum_splits, dim3(grid_for(nw)), dim3(256), 0, n->stream, n->part, nsplit,
                       nw, n->g + l.w_off);
    // bias: column sums of Gb
    const size_t chunk = 4096;
    const int nb = (int)((total + chunk - 1) / chunk);
    rc = ensure_part(n, (size_t)nb * l.cout);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3(nb), dim3(256), 0, n->stream, n->Gb, l.cout, l.cout, total,
                       chunk, n->part);
    hipLaunchKernelGGL(k_sum_splits, dim3(grid_for(l.cout)), dim3(256), 0, n->stream, n->part, nb,
                       (size_t)l.cout, n->g + l.b_off);
    UHIP(n, hipGetLastError());
    return PPF_OK;
}

// Gb = g * leaky'(y) for layer output y (bf16, ldy, yoff) with C channels.
static void act_bwd(ppf_unet *n, const float *g, int ldg, int goff, const bf16 *y, int ldy,
                    int yoff, int C, size_t total)
{
    hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(total * C)), dim3(256), 0, n->stream, g, ldg, goff,
                       y, ldy, yoff, C, total, n->Gb);
}

// Backward through the whole network after forward(); loss into *loss.
static int backward(ppf_unet *n, int B, int D, int H, int W, float *loss)
{
    const int C = n->base;
    const size_t V1 = (size_t)B * D * H * W, V2 = V1 / 8, V3 = V1 / 64;
    const int D2 = D / 2, H2 = H / 2, W2 = W / 2, D3 = D / 4, H3 = H / 4, W3 = W / 4;
    const Layer *L = n->L;
    auto wt = [&](int i) { return n->wt + L[i].wb_off; };
    // head
    const size_t chunk = 4096;
    const int nb = (int)((V1 + chunk - 1) / chunk);
    const int np = 3 * C + 3;
    int rc = ensure_part(n, (size_t)nb * np + nb);
    if (rc)
        return rc;
    float *lossp = n->part + (size_t)nb * np;
    hipLaunchKernelGGL(k_head_bwd, dim3(nb), dim3(256), 0, n->stream, n->yout, n->ytgt, n->D1b, C,
                       n->p + n->head_w, V1, chunk, 1.0f / (float)(V1 * 3), n->gD1b, n->part,
                       lossp);
    hipLaunchKernelGGL(k_sum_splits, dim3(grid_for(np)), dim3(256), 0, n->stream, n->part, nb,
                       (size_t)np, n->g + n->head_w);
    UHIP(n, hipGetLastError());
    std::vector<float> lp(nb);
    UHIP(n, hipMemcpyAsync(lp.data(), lossp, nb * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    // L9: D1a -> D1b
    act_bwd(n, n->gD1b, C, 0, n->D1b, C, 0, C, V1);
    if ((rc = layer_wgrad(n, 9, n->D1a, C, 0, B, D, H, W)))
        return rc;
    conv(n, 1, n->Gb, C, 0, C, wt(9), C, nullptr, C, nullptr, n->gD1a, C, 0, B, D, H, W);
    // L8: U1 -> D1a
    act_bwd(n, n->gD1a, C, 0, n->D1a, C, 0, C, V1);
    if ((rc = layer_wgrad(n, 8, n->U1, 3 * C, 0, B, D, H, W)))
        return rc;
    conv(n, 1, n->Gb, C, 0, C, wt(8), 3 * C, nullptr, 3 * C, nullptr, n->gU1, 3 * C, 0, B, D, H, W);
    // up1
    hipLaunchKernelGGL(k_up_bwd, dim3(grid_for(V2 * 2 * C)), dim3(256), 0, n->stream, n->gU1, 3 * C,
                       0, 2 * C, n->gD2b, B, D, H, W);
    // L7: D2a -> D2b
    act_bwd(n, n->gD2b, 2 * C, 0, n->D2b, 2 * C, 0, 2 * C, V2);
    if ((rc = layer_wgrad(n, 7, n->D2a, 2 * C, 0, B, D2, H2, W2)))
        return rc;
    conv(n, 1, n->Gb, 2 * C, 0, 2 * C, wt(7), 2 * C, nullptr, 2 * C, nullptr, n->gD2a, 2 * C, 0, B,
         D2, H2, W2);
    // L6: U2 -> D2a
    act_bwd(n, n->gD2a, 2 * C, 0, n->D2a, 2 * C, 0, 2 * C, V2);
    if ((rc = layer_wgrad(n, 6, n->U2, 6 * C, 0, B, D2, H2, W2)))
        return rc;
    conv(n, 1, n->Gb, 2 * C, 0, 2 * C, wt(6), 6 * C, nullptr, 6 * C, nullptr, n->gU2, 6 * C, 0, B,
         D2, H2, W2);
    // up2
    hipLaunchKernelGGL(k_up_bwd, dim3(grid_for(V3 * 4 * C)), dim3(256), 0, n->stream, n->gU2, 6 * C,
                       0, 4 * C, n->gE3b, B, D2, H2, W2);
    // L5: E3a -> E3b
    act_bwd(n, n->gE3b, 4 * C, 0, n->E3b, 4 * C, 0, 4 * C, V3);
    if ((rc = layer_wgrad(n, 5, n->E3a, 4 * C, 0, B, D3, H3, W3)))
        return rc;
    conv(n, 1, n->Gb, 4 * C, 0, 4 * C, wt(5), 4 * C, nullptr, 4 * C, nullptr, n->gE3a, 4 * C, 0, B,
         D3, H3, W3);
    // L4: P2 -> E3a
    act_bwd(n, n->gE3a, 4 * C, 0, n->E3a, 4 * C, 0, 4 * C, V3);
    if ((rc = layer_wgrad(n, 4, n->P2, 2 * C, 0, B, D3, H3, W3)))
        return rc;
    conv(n, 1, n->Gb, 4 * C, 0, 4 * C, wt(4), 2 * C, nullptr, 2 * C, nullptr, n->gP2, 2 * C, 0, B,
         D3, H3, W3);
    // pool2: into the skip part of gU2
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for(V2 * 2 * C)), dim3(256), 0, n->stream, n->gP2,
                       2 * C, n->gU2, 6 * C, 4 * C, B, D2, H2, W2);
    // L3: E2a -> E2b (= U2[:, 4C:6C])
    act_bwd(n, n->gU2, 6 * C, 4 * C, n->U2, 6 * C, 4 * C, 2 * C, V2);
    if ((rc = layer_wgrad(n, 3, n->E2a, 2 * C, 0, B, D2, H2, W2)))
        return rc;
    conv(n, 1, n->Gb, 2 * C, 0, 2 * C, wt(3), 2 * C, nullptr, 2 * C, nullptr, n->gE2a, 2 * C, 0, B,
         D2, H2, W2);
    // L2: P1 -> E2a
    act_bwd(n, n->gE2a, 2 * C, 0, n->E2a, 2 * C, 0, 2 * C, V2);
    if ((rc = layer_wgrad(n, 2, n->P1, C, 0, B, D2, H2, W2)))
        return rc;
    conv(n, 1, n->Gb, 2 * C, 0, 2 * C, wt(2), C, nullptr, C, nullptr, n->gP1, C, 0, B, D2, H2, W2);
    // pool1: into the skip part of gU1
    hipLaunchKernelGGL(k_pool_bwd, dim3(grid_for(V1 * C)), dim3(256), 0, n->stream, n->gP1, C,
                       n->gU1, 3 * C, 2 * C, B, D, H, W);
    // L1: E1a -> E1b (= U1[:, 2C:3C])
    act_bwd(n, n->gU1, 3 * C, 2 * C, n->U1, 3 * C, 2 * C, C, V1);
    if ((rc = layer_wgrad(n, 1, n->E1a, C, 0, B, D, H, W)))
        return rc;
    conv(n, 1, n->Gb, C, 0, C, wt(1), C, nullptr, C, nullptr, n->gE1a, C, 0, B, D, H, W);
    // L0: X0 -> E1a (no input gradient needed)
    act_bwd(n, n->gE1a, C, 0, n->E1a, C, 0, C, V1);
    if ((rc = layer_wgrad(n, 0, n->X0, 32, 0, B, D, H, W)))  // X0 is [V][32]
        return rc;
    UHIP(n, hipGetLastError());
    UHIP(n, hipStreamSynchronize(n->stream));
    double s = 0.0;
    for (float v : lp)
        s += v;
    *loss = (float)(s / (double)(V1 * 3));
    return PPF_OK;
}

static bool dims_ok(int B, int D, int H, int W)
{
    return B >= 1 && D >= 4 && H >= 4 && W >= 4 && D % 4 == 0 && H % 4 == 0 && W % 4 == 0 &&
           (size_t)B * D * H * W < ((size_t)1 << 31);
}

extern "C" int ppf_unet_destroy(ppf_unet *n)
{
    if (!n)
        return PPF_OK;
    hipSetDevice(n->device);
    if (n->stream)
        hipStreamSynchronize(n->stream);
    free_ws(n);
    void *bufs[] = {n->p, n->g, n->m, n->v, n->wb, n->wt, n->wc, n->w0, n->wpoly, n->whd, n->tmp, n->part};
    for (void *b : bufs)
        if (b)
            hipFree(b);
    if (n->ev0)
        hipEventDestroy(n->ev0);
    if (n->ev1)
        hipEventDestroy(n->ev1);
    if (n->evc)
        hipEventDestroy(n->evc);
    if (n->evd)
        hipEventDestroy(n->evd);
    if (n->stream)
        hipStreamDestroy(n->stream);
    delete n;
    return PPF_OK;
}

extern "C" int ppf_unet_create(ppf_unet **out, int base, int device, uint64_t seed)
{
    if (!out || (base != 32 && base != 64))
        return uerr(nullptr, PPF_EARG, "base channels must be 32 or 64");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return uerr(nullptr, PPF_ENODEV, "device %d not present", device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return uerr(nullptr, PPF_ENODEV, "device %d is not gfx950", device);
    if (hipSetDevice(device) != hipSuccess)
        return uerr(nullptr, PPF_EHIP, "hipSetDevice failed");
    ppf_unet *n = new ppf_unet();
    n->base = base;
    n->conv_kind = 6;
    n->device = device;
    const int C = base;
    // (cin_log, cin padded, cout) of the ten 3x3x3 convolutions
    const int geo[NCONV][3] = {{3, 32, C},          {C, C, C},         {C, C, 2 * C},
                               {2 * C, 2 * C, 2 * C}, {2 * C, 2 * C, 4 * C}, {4 * C, 4 * C, 4 * C},
                               {6 * C, 6 * C, 2 * C}, {2 * C, 2 * C, 2 * C}, {3 * C, 3 * C, C},
                               {C, C, C}};
    size_t off = 0, woff = 0;
    for (int i = 0; i < NCONV; i++) {
        Layer &l = n->L[i];
        l.cin_log = geo[i][0];
        l.cin = geo[i][1];
        l.cout = geo[i][2];
        l.w_off = off;
        off += (size_t)27 * l.cout * l.cin_log;
        l.b_off = off;
        off += l.cout;
        l.wb_off = woff;
        woff += (size_t)27 * l.cout * l.cin;
    }
    n->head_w = off;
    off += 3 * (size_t)C;
    n->head_b = off;
    off += 3;
    n->np = (int64_t)off;
    n->nwb = woff;
    // polyphase weights of the fused-upsampling layers: L6 over the 4C
    // channels of E3b, L8 over the 2C channels of D2b
    n->wp_off6 = 0;
    n->wp_off8 = (size_t)64 * n->L[6].cout * 4 * C;
    const size_t nwp = n->wp_off8 + (size_t)64 * n->L[8].cout * 2 * C;
    // seeded init: He-uniform for the 3x3x3 layers (gain for leaky ReLU),
    // zero biases, zero head (the untrained network returns its input)
    std::vector<float> h(off, 0.0f);
    uint64_t st = seed;
    auto u01 = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    for (int i = 0; i < NCONV; i++) {
        const Layer &l = n->L[i];
        const double fan_in = 27.0 * l.cin_log;
        const double a = sqrt(6.0 / ((1.0 + LEAK * LEAK) * fan_in));
        for (size_t q = 0; q < (size_t)27 * l.cout * l.cin_log; q++)
            h[l.w_off + q] = (float)((2.0 * u01() - 1.0) * a);
    }
    hipError_t e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&n->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&n->ev1);
    if (e == hipSuccess)
        e = hipMalloc(&n->p, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(&n->g, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(&n->m, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(&n->v, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMalloc(&n->wb, woff * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->wt, woff * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->wc, woff * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->w0, (size_t)C0_CH * C * 32 * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->wpoly, nwp * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->whd, (size_t)C / 32 * 2 * 16 * 32 * sizeof(bf16));
    if (e == hipSuccess)
        e = hipMalloc(&n->tmp, (2 * 1024 + 8) * sizeof(double));
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&n->evc, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipEventCreateWithFlags(&n->evd, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipMemcpy(n->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(n->m, 0, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMemset(n->v, 0, off * sizeof(float));
    if (e == hipSuccess)
        e = hipMemset(n->g, 0, off * sizeof(float));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        uerr(nullptr, PPF_ENOMEM, "network allocation failed: %s", hipGetErrorString(e));
        ppf_unet_destroy(n);
        return PPF_ENOMEM;
    }
    int rc = prep_weights(n);
    if (rc == PPF_OK && hipStreamSynchronize(n->stream) != hipSuccess)
        rc = PPF_EHIP;
    if (rc) {
        uerr(nullptr, rc, "weight preparation failed");
        ppf_unet_destroy(n);
        return rc;
    }
    *out = n;
    return PPF_OK;
}

extern "C" int ppf_unet_nparams(ppf_unet *n, int64_t *np)
{
    if (!n || !np)
        return PPF_EARG;
    *np = n->np;
    return PPF_OK;
}

extern "C" int ppf_unet_get_params(ppf_unet *n, float *params)
{
    if (!n || !params)
        return PPF_EARG;
    UHIP(n, hipSetDevice(n->device));
    UHIP(n, hipStreamSynchronize(n->stream));
    UHIP(n, hipMemcpy(params, n->p, n->np * sizeof(float), hipMemcpyDeviceToHost));
    return PPF_OK;
}

extern "C" int ppf_unet_set_params(ppf_unet *n, const float *params)
{
    if (!n || !params)
        return PPF_EARG;
    UHIP(n, hipSetDevice(n->device));
    UHIP(n, hipStreamSynchronize(n->stream));
    UHIP(n, hipMemcpy(n->p, params, n->np * sizeof(float), hipMemcpyHostToDevice));
    int rc = prep_weights(n);
    if (rc)
        return rc;
    UHIP(n, hipStreamSynchronize(n->stream));
    return PPF_OK;
}

static double forward_flop(const ppf_unet *n, size_t V1)
{
    double f = 0.0;
    const size_t vox[NCONV] = {V1, V1, V1 / 8, V1 / 8, V1 / 64, V1 / 64, V1 / 8, V1 / 8, V1, V1};
    for (int i = 0; i < NCONV; i++)
        f += 2.0 * 27.0 * n->L[i].cin * n->L[i].cout * (double)vox[i];
    if (n->conv_kind >= 2)  // k_conv0: K = 160 instead of 27 x 32
        f -= 2.0 * (27.0 * n->L[0].cin - 32.0 * C0_CH) * n->L[0].cout * (double)vox[0];
    if (n->conv_kind >= 6) {  // polyphase: 8 instead of 27 taps on the upsampled channels
        const int C = n->base;
        f -= 2.0 * (27.0 - 8.0) * (4.0 * C) * n->L[6].cout * (double)vox[6];
        f -= 2.0 * (27.0 - 8.0) * (2.0 * C) * n->L[8].cout * (double)vox[8];
    }
    return f;
}

extern "C" int ppf_unet_forward(ppf_unet *n, int B, int D, int H, int W, const float *x, float *y)
{
    if (!n || !x || !y || !dims_ok(B, D, H, W))
        return n ? uerr(n, PPF_EARG, "bad forward arguments") : PPF_EARG;
    UHIP(n, hipSetDevice(n->device));
    int rc = ensure_ws(n, B, D, H, W, false);
    if (rc)
        return rc;
    const size_t V1 = (size_t)B * D * H * W;
    UHIP(n, hipMemcpyAsync(n->xin, x, V1 * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    UHIP(n, hipEventRecord(n->ev0, n->stream));
    if ((rc = forward(n, B, D, H, W, false)))
        return rc;
    UHIP(n, hipEventRecord(n->ev1, n->stream));
    UHIP(n, hipMemcpyAsync(y, n->yout, V1 * 3 * sizeof(float), hipMemcpyDeviceToHost, n->stream));
    UHIP(n, hipStreamSynchronize(n->stream));
    float ms = 0.f;
    UHIP(n, hipEventElapsedTime(&ms, n->ev0, n->ev1));
    n->last_ms = ms;
    n->timing_pending = false;
    n->last_flop = forward_flop(n, V1);
    return PPF_OK;
}

static int loss_grad(ppf_unet *n, int B, int D, int H, int W, const float *x, const float *y,
                     float *loss)
{
    int rc = ensure_ws(n, B, D, H, W, true);
    if (rc)
        return rc;
    const size_t V1 = (size_t)B * D * H * W;
    UHIP(n, hipMemcpyAsync(n->xin, x, V1 * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    UHIP(n, hipMemcpyAsync(n->ytgt, y, V1 * 3 * sizeof(float), hipMemcpyHostToDevice, n->stream));
    if ((rc = forward(n, B, D, H, W, true)))
        return rc;
    return backward(n, B, D, H, W, loss);
}

extern "C" int ppf_unet_loss_grad(ppf_unet *n, int B, int D, int H, int W, const float *x,
                                  const float *y, float *loss, float *grads)
{
    if (!n || !x || !y || !loss || !dims_ok(B, D, H, W))
        return n ? uerr(n, PPF_EARG, "bad arguments") : PPF_EARG;
    UHIP(n, hipSetDevice(n->device));
    int rc = loss_grad(n, B, D, H, W, x, y, loss);
    if (rc)
        return rc;
    if (grads)
        UHIP(n, hipMemcpy(grads, n->g, n->np * sizeof(float), hipMemcpyDeviceToHost));
    return PPF_OK;
}

extern "C" int ppf_unet_train_step(ppf_unet *n, int B, int D, int H, int W, const float *x,
                                   const float *y, float lr, float *loss)
{
    if (!n || !x || !y || !loss || !dims_ok(B, D, H, W) || !(lr >= 0.f))
        return n ? uerr(n, PPF_EARG, "bad arguments") : PPF_EARG;
    UHIP(n, hipSetDevice(n->device));
    int rc = loss_grad(n, B, D, H, W, x, y, loss);
    if (rc)
        return rc;
    n->t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)n->t), c2 = 1.f - powf(b2, (float)n->t);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(n->np)), dim3(256), 0, n->stream, n->p, n->g, n->m,
                       n->v, (size_t)n->np, lr, b1, b2, 1e-8f, c1, c2);
    UHIP(n, hipGetLastError());
    if ((rc = prep_weights(n)))
        return rc;
    UHIP(n, hipStreamSynchronize(n->stream));
    return PPF_OK;
}

extern "C" int ppf_unet_apply(ppf_unet *n, ppf_ctx *ctx)
{
    if (!n || !ctx)
        return n ? uerr(n, PPF_EARG, "null context") : PPF_EARG;
    const int D = ctx->d.nz, H = ctx->d.ny, W = ctx->d.nx;
    if (!dims_ok(1, D, H, W))
        return uerr(n, PPF_EARG, "grid %dx%dx%d: every dimension must be a multiple of 4", D, H,
                    W);
    if (ctx->d.device != n->device)
        return uerr(n, PPF_EARG, "network and context are on different devices");
    UHIP(n, hipSetDevice(n->device));
    int rc = ensure_ws(n, 1, D, H, W, false);
    if (rc)
        return rc;
    // stream order: the context's queued steps, then this, then the
    // context's later work (events; no host synchronisation)
    UHIP(n, hipEventRecord(n->evc, ctx->stream));
    UHIP(n, hipStreamWaitEvent(n->stream, n->evc, 0));
    const size_t N = ctx->n;
    void *fs = ctx->buf[ctx->cur[0]][0], *fl = ctx->buf[ctx->cur[1]][1],
         *fc = ctx->buf[ctx->cur[2]][2];
    const bool f64 = ctx->esz == 8;
    double *tmp = n->tmp;
    // solute before the jump (FP64, double-double): tmp[2049]
    if (f64)
        hipLaunchKernelGGL(k_sum_dd<double>, dim3(1024), dim3(256), 0, n->stream,
                           (const double *)fc, N, tmp);
    else
        hipLaunchKernelGGL(k_sum_dd<float>, dim3(1024), dim3(256), 0, n->stream,
                           (const float *)fc, N, tmp);
    hipLaunchKernelGGL(k_mass_delta, dim3(1), dim3(256), 0, n->stream, tmp, 1024, 0.0, 1.0,
                       tmp + 2048);
    // the row-blocked kernels read and write the context's fields directly;
    // the others go through the packed FP32 copies
    const bool direct = n->conv_kind >= 2;
    const CtxIO io{fs, fl, fc, f64};
    if (!direct) {
        if (f64)
            hipLaunchKernelGGL(k_ctx_to_x<double>, dim3(grid_for(N)), dim3(256), 0, n->stream,
                               (const double *)fs, (const double *)fl, (const double *)fc, N,
                               n->xin);
        else
            hipLaunchKernelGGL(k_ctx_to_x<float>, dim3(grid_for(N)), dim3(256), 0, n->stream,
                               (const float *)fs, (const float *)fl, (const float *)fc, N, n->xin);
    }
    UHIP(n, hipEventRecord(n->ev0, n->stream));
    rc = forward(n, 1, D, H, W, false, direct ? &io : nullptr);
    UHIP(n, hipEventRecord(n->ev1, n->stream));
    if (rc)
        return rc;
    if (!direct) {
        if (f64)
            hipLaunchKernelGGL(k_y_to_ctx<double>, dim3(grid_for(N)), dim3(256), 0, n->stream,
                               n->yout, N, (double *)fs, (double *)fl, (double *)fc);
        else
            hipLaunchKernelGGL(k_y_to_ctx<float>, dim3(grid_for(N)), dim3(256), 0, n->stream,
                               n->yout, N, (float *)fs, (float *)fl, (float *)fc);
    }
    // project c back onto the pre-jump total: delta = (before - after) / N
    // into tmp[2052], then the shift (FP32: stochastically rounded)
    if (f64) {
        hipLaunchKernelGGL(k_sum_dd<double>, dim3(1024), dim3(256), 0, n->stream,
                           (const double *)fc, N, tmp);
        hipLaunchKernelGGL(k_mass_delta_dev, dim3(1), dim3(256), 0, n->stream, tmp, 1024,
                           (const double *)(tmp + 2049), (double)N, tmp + 2052);
        hipLaunchKernelGGL(k_add_delta, dim3(grid_for(N)), dim3(256), 0, n->stream, (double *)fc,
                           N, (const double *)(tmp + 2052));
    } else {
        hipLaunchKernelGGL(k_sum_dd<float>, dim3(1024), dim3(256), 0, n->stream,
                           (const float *)fc, N, tmp);
        hipLaunchKernelGGL(k_mass_delta_dev, dim3(1), dim3(256), 0, n->stream, tmp, 1024,
                           (const double *)(tmp + 2049), (double)N, tmp + 2052);
        hipLaunchKernelGGL(k_mass_shift, dim3(grid_for(N)), dim3(256), 0, n->stream, (float *)fc,
                           N, (const double *)(tmp + 2052), (uint64_t)ctx->steps);
    }
    UHIP(n, hipGetLastError());
    UHIP(n, hipEventRecord(n->evd, n->stream));
    UHIP(n, hipStreamWaitEvent(ctx->stream, n->evd, 0));
    n->timing_pending = true;
    n->last_flop = forward_flop(n, N);
    return PPF_OK;
}

extern "C" int ppf_unet_last_forward(ppf_unet *n, double *ms, double *flop)
{
    if (!n || !ms || !flop)
        return PPF_EARG;
    if (n->timing_pending) {  // an apply's forward: read its events now
        UHIP(n, hipEventSynchronize(n->ev1));
        float f = 0.f;
        UHIP(n, hipEventElapsedTime(&f, n->ev0, n->ev1));
        n->last_ms = f;
        n->timing_pending = false;
    }
    *ms = n->last_ms;
    *flop = n->last_flop;
    return PPF_OK;
}

extern "C" int ppf_unet_set_conv_kernel(ppf_unet *n, int kind)
{
    if (!n || kind < 0 || kind > 7)
        return n ? uerr(n, PPF_EARG, "convolution kernel must be 0..7") : PPF_EARG;
    n->conv_kind = kind;
    return PPF_OK;
}

extern "C" const char *ppf_unet_last_error(ppf_unet *n) { return n ? n->err : g_unet_err; }
