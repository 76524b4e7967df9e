// ppf.hip -- libppf.so: context, launches and the C ABI of include/ppf.h.
//
// The fields live in HBM for the life of a context as two ping-pong copies
// of (etaS, etaL, c), dense [nz][ny][nx] with x fastest.  One step is one
// kernel launch on the context's stream; the ghost fill of the reference
// (phasefield.c:1112-1189) is the kernel's index map.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ppf.h"
#include "ppf_host.h"
#include "k_step.h"
#include "k_misc.h"
#include "k_ccl.h"
#include "init_spheres.h"

#include <algorithm>
#include <mutex>
#include <type_traits>
#include <vector>

constexpr int MAX_DIM = 1 << 16;

// Message of the last failure without a context (ppf_create, the no-alloc
// sweep), one per host thread.
static thread_local char g_create_err[512];

static int set_err(ppf_ctx *ctx, int code, const char *fmt, ...)
{
    char *dst = ctx ? ctx->err : g_create_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, 512, fmt, ap);
    va_end(ap);
    return code;
}

// A failed HIP call also leaves its error as the thread's "last error",
// which a later, unrelated hipGetLastError would report; every error path
// clears it (hipGetLastError resets it) so that a context or call after an
// error (an out-of-memory create, say) starts clean.
#define HIPCHK(ctx, call)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            (void)hipGetLastError();                                                        \
            return set_err((ctx), PPF_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
        }                                                                                   \
    } while (0)

extern "C" int ppf_params_default(ppf_params *p)
{
    if (!p)
        return PPF_EARG;
    // phasefield.c:76-97, same literals and expression order.
    const double AS = 1e+09;
    const double AL = 1e+09;
    const double CS = -1e+06;
    const double CL = 1.e+07;
    const double cS0 = 0.455;
    const double cL0 = 0.978;
    const double cSeq = 0.999;
    const double cLeq = 0.476;
    const double DL = 1.0e-14, DS = 0.0;
    const double ML = DL / AL, MS = DS / AS;
    const double sigma = 2.0;
    const double gamma = 1.5;
    const double dx = 1;
    const double h2 = 1.0 / (dx * dx);
    const double dt = 0.125 * dx * dx / DL;
    const double l = 7.0 * dx;
    const double m = 6.0 * sigma / l;
    const double kappa = 3 * 0.25 * sigma * l;
    const double L = 4.0 / 3.0 * m / kappa * (0.5 * (MS + ML) / ((cSeq - cLeq) * (cSeq - cLeq)));
    p->MS = MS;
    p->ML = ML;
    p->AS = AS;
    p->AL = AL;
    p->cL0 = cL0;
    p->cS0 = cS0;
    p->CS = CS;
    p->CL = CL;
    p->L = L;
    p->dt = dt;
    p->h2 = h2;
    p->m = m;
    p->gamma = gamma;
    p->kappa = kappa;
    return PPF_OK;
}

// ----------------------------------------------------------------------------
// Launch tables
// ----------------------------------------------------------------------------
template <class P>
using tiled_fn = void (*)(const typename P::T *, const typename P::T *, const typename P::T *,
                          typename P::T *, typename P::T *, typename P::T *, int, int, int, int,
                          int, int, int, typename P::K, double *);
typedef void (*naive_fn)(const double *, const double *, const double *, double *, double *,
                         double *, int, int, int, KConst, double *);

template <class P, int TX, int TY, bool PIPE, int CD = 2>
static tiled_fn<P> pick_tiled_p(int bc, bool conly, bool report)
{
#define PPF_T(BC, C, R) \
    if (bc == BC && conly == C && report == R) return k_tiled<P, TX, TY, BC, C, R, PIPE, CD>;
    PPF_T(0, false, false) PPF_T(0, false, true) PPF_T(0, true, false) PPF_T(0, true, true)
    PPF_T(1, false, false) PPF_T(1, false, true) PPF_T(1, true, false) PPF_T(1, true, true)
#undef PPF_T
    return nullptr;
}

// order bit 2 selects the unpipelined march; bit 3 (FP64, pipelined) the
// column loaded four planes ahead instead of two
template <class P, int TX, int TY>
static tiled_fn<P> pick_tiled(int bc, bool conly, bool report, int order)
{
    if (order & 4)
        return pick_tiled_p<P, TX, TY, false>(bc, conly, report);
    if constexpr (sizeof(typename P::T) == 8)
        if (order & 8)
            return pick_tiled_p<P, TX, TY, true, 4>(bc, conly, report);
    return pick_tiled_p<P, TX, TY, true>(bc, conly, report);
}

static naive_fn pick_naive(int bc, bool eqa, bool conly, bool report)
{
#define PPF_N(BC, E, C, R) \
    if (bc == BC && eqa == E && conly == C && report == R) return k_naive_f64<BC, E, C, R>;
#define PPF_N4(BC, E) PPF_N(BC, E, false, false) PPF_N(BC, E, false, true) \
                      PPF_N(BC, E, true, false) PPF_N(BC, E, true, true)
    PPF_N4(0, true) PPF_N4(0, false) PPF_N4(1, true) PPF_N4(1, false)
#undef PPF_N4
#undef PPF_N
    return nullptr;
}

// Tiled contexts run one physics policy: PhysF32 (FP32 form), PhysF64F
// (the fast FP64 form, PPF_KERNEL_FAST), PhysF64<true> (bitwise, AS == AL)
// or PhysF64<false> (bitwise, general constants).  fn(policy tag,
// constants) is called with the context's.
template <class Fn>
static int with_policy(const ppf_ctx *ctx, Fn &&fn)
{
    if (ctx->d.precision == PPF_F32)
        return fn(PhysF32{}, ctx->k32);
    if (ctx->kernel == PPF_KERNEL_FAST)
        return fn(PhysF64F{}, ctx->kf);
    if (ctx->eqa)
        return fn(PhysF64<true>{}, ctx->k);
    return fn(PhysF64<false>{}, ctx->k);
}

static bool tiled_kernel(const ppf_ctx *ctx)
{
    return ctx->kernel == PPF_KERNEL_TILED || ctx->kernel == PPF_KERNEL_FAST;
}

// Column-tile shapes compiled for the tiled kernel (TX x TY).
static const TileCfg k_tiles[] = {{64, 8}, {64, 16}, {64, 4}, {32, 16}, {32, 8}, {16, 16}};

template <class P>
static tiled_fn<P> pick_tiled_shape(TileCfg t, int bc, bool conly, bool report, int order)
{
    if (t.tx == 64 && t.ty == 8)
        return pick_tiled<P, 64, 8>(bc, conly, report, order);
    if (t.tx == 64 && t.ty == 16)
        return pick_tiled<P, 64, 16>(bc, conly, report, order);
    if (t.tx == 64 && t.ty == 4)
        return pick_tiled<P, 64, 4>(bc, conly, report, order);
    if (t.tx == 32 && t.ty == 16)
        return pick_tiled<P, 32, 16>(bc, conly, report, order);
    if (t.tx == 32 && t.ty == 8)
        return pick_tiled<P, 32, 8>(bc, conly, report, order);
    if (t.tx == 16 && t.ty == 16)
        return pick_tiled<P, 16, 16>(bc, conly, report, order);
    return nullptr;
}

static void tiled_grid(const ppf_ctx *ctx, int &ntx, int &nty, int &nzc)
{
    ntx = (ctx->d.nx + ctx->tile.tx - 1) / ctx->tile.tx;
    nty = (ctx->d.ny + ctx->tile.ty - 1) / ctx->tile.ty;
    nzc = (ctx->d.nz + ctx->tz - 1) / ctx->tz;
}

static size_t report_blocks(const ppf_ctx *ctx)
{
    if (ctx->kernel == PPF_KERNEL_NAIVE)
        return (size_t)((ctx->d.nx + 63) / 64) * ((ctx->d.ny + 3) / 4) * ctx->d.nz;
    int ntx, nty, nzc;
    tiled_grid(ctx, ntx, nty, nzc);
    return (size_t)ntx * nty * nzc;
}

static int ensure_partials(ppf_ctx *ctx, size_t ndoubles)
{
    if (ndoubles <= ctx->partials_cap)
        return PPF_OK;
    if (ctx->partials)
        HIPCHK(ctx, hipFree(ctx->partials));
    ctx->partials = nullptr;
    ctx->partials_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->partials, ndoubles * sizeof(double)));
    ctx->partials_cap = ndoubles;
    return PPF_OK;
}

template <class P>
static int launch_tiled(ppf_ctx *ctx, const typename P::K &k, bool conly, double *part)
{
    using T = typename P::T;
    const int cs = ctx->cur[0], cl = ctx->cur[1], cc = ctx->cur[2];
    int ntx, nty, nzc;
    tiled_grid(ctx, ntx, nty, nzc);
    tiled_fn<P> f =
        pick_tiled_shape<P>(ctx->tile, ctx->d.bc, conly, part != nullptr, ctx->order);
    if (!f)
        return set_err(ctx, PPF_EARG, "no kernel for tile %dx%d", ctx->tile.tx, ctx->tile.ty);
    hipLaunchKernelGGL(f, dim3((unsigned)((size_t)ntx * nty * nzc)),
                       dim3(ctx->tile.tx * ctx->tile.ty), 0, ctx->stream,
                       (const T *)ctx->buf[cs][0], (const T *)ctx->buf[cl][1],
                       (const T *)ctx->buf[cc][2], (T *)ctx->buf[1 - cs][0],
                       (T *)ctx->buf[1 - cl][1], (T *)ctx->buf[1 - cc][2], ctx->d.nx, ctx->d.ny,
                       ctx->d.nz, ctx->tz, ntx, nty, ctx->order, k, part);
    return PPF_OK;
}

// Two steps per launch (k_tiled2, k_tiled2s).
template <class P>
using fused_fn = void (*)(const typename P::T *, const typename P::T *, const typename P::T *,
                          typename P::T *, typename P::T *, typename P::T *, int, int, int, int,
                          int, int, int, typename P::K);

static const TileCfg k_tiles2[] = {{64, 8}, {32, 8}, {32, 16}, {64, 16}, {16, 16}};

template <class P>
struct FusedPick {
    fused_fn<P> f;
    int threads;
};

template <class P>
static FusedPick<P> pick_fused(TileCfg t, int order, int bc = PPF_BC_PERIODIC, bool full = false)
{
    // k_tiled2: FP64 needs ~176 VGPRs: two waves per SIMD; at two waves per
    // SIMD the loads must be issued a plane ahead (PF) to be hidden.  Order
    // bit 2 (FP64, 4-wave tiles 32x8 and 16x16): three waves per SIMD
    // instead, without PF (164 VGPRs, no spills; 49 KB of LDS, so three
    // workgroups share a CU).  FP32 fits five waves.  LDS double buffered
    // (one barrier per plane) except for the FP64 64x16 tile, whose two
    // buffers would not fit.
    // Order bit 3: k_tiled2s, the ring sites in waves of their own after the
    // tile's (FP64 114 VGPRs: four waves per SIMD; FP32 55: eight).
    // full: the tile divides the plane; the rotating march (the fast FP64
    // and the FP32 forms, no early loads) then defers its output stores.
    constexpr bool F64 = sizeof(typename P::T) == 8;
    constexpr int W = F64 ? 2 : 5, WS = F64 ? 4 : 8;
    const int nt = t.tx * t.ty;
    if (bc == PPF_BC_NOFLUX) {
        // mirror boundaries: FP64 at three waves per SIMD, FP32 at five,
        // double-buffered LDS (order bits 2 and 3 have no effect)
        constexpr int WN = F64 ? 3 : 5;
        if (t.tx == 16 && t.ty == 16)
            return {k_tiled2<P, 16, 16, WN, false, true, PPF_BC_NOFLUX>, nt};
        if (t.tx == 32 && t.ty == 8)
            return {k_tiled2<P, 32, 8, WN, false, true, PPF_BC_NOFLUX>, nt};
        if (!F64 && t.tx == 32 && t.ty == 16)
            return {k_tiled2<P, 32, 16, WN, false, true, PPF_BC_NOFLUX>, nt};
        return {nullptr, 0};
    }
    if (order & 8) {
        const int ring = (2 * t.tx + 2 * t.ty + 63) / 64 * 64;
        if (t.tx == 16 && t.ty == 16)
            return {k_tiled2s<P, 16, 16, WS>, nt + ring};
        if (t.tx == 32 && t.ty == 8)
            return {k_tiled2s<P, 32, 8, WS>, nt + ring};
        if (t.tx == 32 && t.ty == 16)
            return {k_tiled2s<P, 32, 16, WS>, nt + ring};
        return {nullptr, 0};
    }
    const bool w3 = F64 && (order & 4);
    // Order bit 4: loads issued a plane early (PF) in the variants whose
    // registers allow it: the fast FP64 form at three waves per SIMD (160
    // VGPRs; the bitwise form would spill) and FP32 at five.
    constexpr bool PF_OK = !F64 || std::is_same<P, PhysF64F>::value;
    if constexpr (PF_OK) {
        if (order & 16) {
            if (F64 && !w3)
                return {nullptr, 0};
            constexpr int WP = F64 ? 3 : 5;
            if (t.tx == 32 && t.ty == 8)
                return {k_tiled2<P, 32, 8, WP, true, true>, nt};
            if (t.tx == 16 && t.ty == 16)
                return {k_tiled2<P, 16, 16, WP, true, true>, nt};
            if (!F64 && t.tx == 32 && t.ty == 16)
                return {k_tiled2<P, 32, 16, WP, true, true>, nt};
            if (!F64 && t.tx == 64 && t.ty == 8)
                return {k_tiled2<P, 64, 8, WP, true, true>, nt};
            return {nullptr, 0};
        }
    } else if (order & 16) {
        return {nullptr, 0};
    }
    constexpr bool ROT = P::kRotate;
    if (t.tx == 64 && t.ty == 8) {
        if constexpr (!F64 && ROT) {
            if (full)
                return {k_tiled2<P, 64, 8, W, false, true, 0, true>, nt};
        }
        return {k_tiled2<P, 64, 8, W, F64, true>, nt};
    }
    if (t.tx == 32 && t.ty == 8) {
        if constexpr (ROT) {
            if (full && (w3 || !F64))
                return {k_tiled2<P, 32, 8, F64 ? 3 : W, false, true, 0, true>, nt};
        }
        return {w3 ? k_tiled2<P, 32, 8, 3, false, true> : k_tiled2<P, 32, 8, W, F64, true>, nt};
    }
    if (t.tx == 16 && t.ty == 16) {
        if constexpr (ROT) {
            if (full && (w3 || !F64))
                return {k_tiled2<P, 16, 16, F64 ? 3 : W, false, true, 0, true>, nt};
        }
        return {w3 ? k_tiled2<P, 16, 16, 3, false, true> : k_tiled2<P, 16, 16, W, F64, true>, nt};
    }
    if (t.tx == 32 && t.ty == 16) {
        if constexpr (!F64 && ROT) {
            if (full)
                return {k_tiled2<P, 32, 16, W, false, true, 0, true>, nt};
        }
        return {k_tiled2<P, 32, 16, W, F64, true>, nt};
    }
    if (t.tx == 64 && t.ty == 16)
        return {k_tiled2<P, 64, 16, W == 2 ? 1 : W, F64, !F64>, nt};
    return {nullptr, 0};
}

// Whether a two-step kernel is compiled for tile t and order o at the
// context's precision and boundary condition.
static bool fused_exists(const ppf_ctx *ctx, TileCfg t, int o)
{
    return with_policy(ctx, [&](auto p, const auto &) {
               return (int)(pick_fused<decltype(p)>(t, o, ctx->d.bc).f != nullptr);
           }) != 0;
}

static bool fused_available(const ppf_ctx *ctx)
{
    return ctx->fuse == 2 && tiled_kernel(ctx);
}

template <class P>
static int launch_fused(ppf_ctx *ctx, const typename P::K &k)
{
    using T = typename P::T;
    const int cs = ctx->cur[0], cl = ctx->cur[1], cc = ctx->cur[2];
    const int ntx = (ctx->d.nx + ctx->tile2.tx - 1) / ctx->tile2.tx;
    const int nty = (ctx->d.ny + ctx->tile2.ty - 1) / ctx->tile2.ty;
    const int nzc = (ctx->d.nz + ctx->tz2 - 1) / ctx->tz2;
    const FusedPick<P> pk =
        pick_fused<P>(ctx->tile2, ctx->order2, ctx->d.bc,
                      ctx->d.nx % ctx->tile2.tx == 0 && ctx->d.ny % ctx->tile2.ty == 0);
    if (!pk.f)
        return set_err(ctx, PPF_EARG, "no two-step kernel for tile %dx%d order %d", ctx->tile2.tx,
                       ctx->tile2.ty, ctx->order2);
    hipLaunchKernelGGL(pk.f, dim3((unsigned)((size_t)ntx * nty * nzc)), dim3(pk.threads), 0,
                       ctx->stream,
                       (const T *)ctx->buf[cs][0], (const T *)ctx->buf[cl][1],
                       (const T *)ctx->buf[cc][2], (T *)ctx->buf[1 - cs][0],
                       (T *)ctx->buf[1 - cl][1], (T *)ctx->buf[1 - cc][2], ctx->d.nx, ctx->d.ny,
                       ctx->d.nz, ctx->tz2, ntx, nty, ctx->order2, k);
    return PPF_OK;
}

// Time `reps` sweeps of the current state without advancing it: the kernel
// reads copy cur and writes copy 1-cur, and cur is left as it was.
static int time_tiling(ppf_ctx *ctx, int reps, double *ms)
{
    int rc = PPF_OK;
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps && rc == PPF_OK; r++)
        rc = with_policy(ctx, [&](auto p, const auto &k) {
            return launch_tiled<decltype(p)>(ctx, k, false, nullptr);
        });
    if (rc)
        return rc;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = (double)f / reps;
    return PPF_OK;
}

// Launch one sweep from copy cur to copy 1-cur; swaps the moved fields.
static int launch_step(ppf_ctx *ctx, bool conly, bool report)
{
    const int nx = ctx->d.nx, ny = ctx->d.ny, nz = ctx->d.nz;
    const int bc = ctx->d.bc;
    double *part = report ? ctx->partials : nullptr;
    const int cs = ctx->cur[0], cl = ctx->cur[1], cc = ctx->cur[2];
    int rc = PPF_OK;
    if (ctx->kernel == PPF_KERNEL_NAIVE) {
        naive_fn f = pick_naive(bc, ctx->eqa, conly, report);
        dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
        hipLaunchKernelGGL(f, grid, dim3(256), 0, ctx->stream, (const double *)ctx->buf[cs][0],
                           (const double *)ctx->buf[cl][1], (const double *)ctx->buf[cc][2],
                           (double *)ctx->buf[1 - cs][0], (double *)ctx->buf[1 - cl][1],
                           (double *)ctx->buf[1 - cc][2], nx, ny, nz, ctx->k, part);
    } else {
        rc = with_policy(ctx, [&](auto p, const auto &k) {
            return launch_tiled<decltype(p)>(ctx, k, conly, part);
        });
    }
    if (rc)
        return rc;
    HIPCHK(ctx, hipGetLastError());
    if (!conly) {
        ctx->cur[0] ^= 1;
        ctx->cur[1] ^= 1;
    }
    ctx->cur[2] ^= 1;
    ctx->steps++;
    return PPF_OK;
}

// Two sweeps in one launch (fused_available), from copy cur to 1-cur.
static int launch_step2(ppf_ctx *ctx)
{
    const int rc = with_policy(
        ctx, [&](auto p, const auto &k) { return launch_fused<decltype(p)>(ctx, k); });
    if (rc)
        return rc;
    HIPCHK(ctx, hipGetLastError());
    ctx->cur[0] ^= 1;
    ctx->cur[1] ^= 1;
    ctx->cur[2] ^= 1;
    ctx->steps += 2;
    return PPF_OK;
}

// FP32 solute bookkeeping (no-op for FP64): after a step, every shift_every
// steps, sum c in FP64 on the device and shift c back onto the FP64 target
// with stochastic rounding (k_mass_shift), all on the context stream.
static int after_step(ppf_ctx *ctx)
{
    if (ctx->d.precision != PPF_F32 || ctx->shift_every <= 0)
        return PPF_OK;
    if (++ctx->since_shift < ctx->shift_every)
        return PPF_OK;
    ctx->since_shift = 0;
    const int nb = 1024;
    int rc = ensure_partials(ctx, (size_t)2 * nb);
    if (rc)
        return rc;
    float *c = (float *)ctx->buf[ctx->cur[2]][2];
    hipLaunchKernelGGL(k_sum_dd<float>, dim3(nb), dim3(256), 0, ctx->stream, (const float *)c,
                       ctx->n, ctx->partials);
    hipLaunchKernelGGL(k_mass_delta, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, nb,
                       ctx->c_target, (double)ctx->n, ctx->dres + 8);
    const unsigned g = (unsigned)((ctx->n + 255) / 256 < 65536 ? (ctx->n + 255) / 256 : 65536);
    hipLaunchKernelGGL(k_mass_shift, dim3(g), dim3(256), 0, ctx->stream, c, ctx->n,
                       (const double *)(ctx->dres + 8), (uint64_t)ctx->steps);
    HIPCHK(ctx, hipGetLastError());
    ctx->n_shifts++;
    return PPF_OK;
}

// FP64 target of sum(c) from the device field (FP32 contexts).
static int set_target_from_device(ppf_ctx *ctx)
{
    if (ctx->d.precision != PPF_F32)
        return PPF_OK;
    const int nb = 1024;
    int rc = ensure_partials(ctx, (size_t)2 * nb);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum_dd<float>, dim3(nb), dim3(256), 0, ctx->stream,
                       (const float *)ctx->buf[ctx->cur[2]][2], ctx->n, ctx->partials);
    hipLaunchKernelGGL(k_mass_delta, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, nb, 0.0,
                       1.0, ctx->dres + 8);
    HIPCHK(ctx, hipGetLastError());
    double r[2];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->dres + 8, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->c_target = r[1];
    ctx->since_shift = 0;
    return PPF_OK;
}

// Default tiling, from sweeps on MI355X at 512^3: single steps in x rows
// of 64 sites and 32-plane z chunks; FP64 pipelined 64x8 tiles in
// tile-fastest order, FP32 unpipelined 64x16 tiles with the z chunks of a
// tile adjacent.
// ppf_set_tiling or ppf_autotune override it.
static void choose_tiles(ppf_ctx *ctx)
{
    const int nx = ctx->d.nx;
    const bool f32 = ctx->d.precision == PPF_F32;
    if (nx >= 64)
        ctx->tile = f32 ? TileCfg{64, 16} : TileCfg{64, 8};
    else if (nx >= 32)
        ctx->tile = {32, 8};
    else
        ctx->tile = {16, 16};
    ctx->order = f32 ? 5 : 0;
    ctx->tz = ctx->d.nz < 32 ? ctx->d.nz : 32;
    // two steps per launch: on by default (periodic and mirror boundaries;
    // the mirror kernels have no ring waves and no order bits 2-3).  FP32: 32x16
    // tiles, tz 64, XCD order (1.2x the single-step rate at 512^3; level
    // with 32x8 there and 14% faster at 256^3, 233 against 205 GLUPS).
    // FP64: 16x16 tiles, XCD order, three waves per SIMD (order bit 2), the
    // z chunk chosen below: 117-121 against 101-103 GLUPS for single steps
    // at 512^3, 114-117 against 93-98 at 256^3, 116-118 against 87-93 at
    // 1024^3.
    // FP32 runs the ring-wave kernel (order bit 3): 221 against 211 GLUPS at
    // 256^3, level at 512^3.
    // The fast FP64 form: 32x8 tiles (145 against 142 GLUPS for 16x16 at
    // 512^3, 130-134 against 119 at 256^3, level at 1024^3).  FP32 on grids
    // of 512^3 sites and more, and from 256^3 on where 64x8 tiles divide
    // the plane (the deferred-store form): 64x8 tiles without ring waves
    // (order 2; 305 against 258 GLUPS for the ring-wave kernel at 512^3, 299
    // against 263-273 at 256^3); otherwise the ring waves stay.
    ctx->fuse = 2;
    const bool big = ctx->n >= ((size_t)1 << 27) ||
                     (ctx->n >= ((size_t)1 << 24) && ctx->d.nx % 64 == 0 && ctx->d.ny % 8 == 0);
    ctx->tile2 = f32 ? (big ? TileCfg{64, 8} : TileCfg{32, 16})
                     : ctx->kernel == PPF_KERNEL_FAST ? TileCfg{32, 8} : TileCfg{16, 16};
    ctx->order2 = f32 ? (big ? 2 : 10) : 6;
    // z chunks of 64 planes, shortened (to 16 at least) while the launch
    // would leave CUs without a workgroup: at 128^3, 64-plane chunks of
    // 32x16 tiles are 64 workgroups (38 GLUPS), 32-plane ones 128 (75).
    const long tiles2 = (long)((ctx->d.nx + ctx->tile2.tx - 1) / ctx->tile2.tx) *
                        ((ctx->d.ny + ctx->tile2.ty - 1) / ctx->tile2.ty);
    int tz2 = 64;
    while (tz2 > 16 && tiles2 * ((ctx->d.nz + tz2 - 1) / tz2) < 256)
        tz2 /= 2;
    // FP64: the z chunk whose workgroups fill whole rounds of the resident
    // slots (CUs x workgroups per CU, three for the 16x16 three-wave
    // kernel), longer chunks preferred (less overlap): 86 planes at 256^3
    // (768 workgroups on 256 CUs, 114.7 against 105.6 GLUPS with 64-plane
    // chunks) and at 512^3 (118.9 against 118.0).  FP32 keeps 64 (best at
    // 256^3 and 512^3).
    if (!f32) {
        int ncu = 0, per_cu = 0;
        const void *kf = ctx->kernel == PPF_KERNEL_FAST
                             ? reinterpret_cast<const void *>(
                                   pick_fused<PhysF64F>(ctx->tile2, ctx->order2).f)
                             : reinterpret_cast<const void *>(
                                   pick_fused<PhysF64<true>>(ctx->tile2, ctx->order2).f);
        if (kf && hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                        ctx->d.device) == hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf,
                                                         ctx->tile2.tx * ctx->tile2.ty,
                                                         0) == hipSuccess &&
            ncu > 0 && per_cu > 0) {
            const double slots = (double)ncu * per_cu;
            double best = -1.0;
            for (int c : {128, 86, 64, 43, 32, 16}) {
                const double wgs = (double)tiles2 * ((ctx->d.nz + c - 1) / c);
                const double rounds = ceil(wgs / slots);
                const double score = wgs / (rounds * slots) * c / (c + 2.0);
                if (score > best + 1e-9) {
                    best = score;
                    tz2 = c;
                }
            }
        }
    }
    ctx->tz2 = ctx->d.nz < tz2 ? ctx->d.nz : tz2;
}

extern "C" int ppf_create(ppf_ctx **out, const ppf_desc *desc)
{
    if (!out || !desc)
        return set_err(nullptr, PPF_EARG, "null argument");
    *out = nullptr;
    const ppf_desc &d = *desc;
    if (d.nz < 1 || d.ny < 1 || d.nx < 1 || d.nz > MAX_DIM || d.ny > MAX_DIM || d.nx > MAX_DIM)
        return set_err(nullptr, PPF_EARG, "bad grid %d x %d x %d", d.nz, d.ny, d.nx);
    if (d.bc != PPF_BC_PERIODIC && d.bc != PPF_BC_NOFLUX)
        return set_err(nullptr, PPF_EARG, "bad bc %d", d.bc);
    if (d.precision != PPF_F64 && d.precision != PPF_F32)
        return set_err(nullptr, PPF_EARG, "bad precision %d", d.precision);
    if (d.precision == PPF_F32 && d.params.AS != d.params.AL)
        return set_err(nullptr, PPF_EARG, "the FP32 form needs AS == AL");
    if (d.precision == PPF_F32 && !(d.params.ML > 0.0))
        return set_err(nullptr, PPF_EARG, "the FP32 form needs ML > 0");
    if (d.precision == PPF_F32 && d.kernel == PPF_KERNEL_NAIVE)
        return set_err(nullptr, PPF_EARG, "the naive kernel is FP64 only");
    if (d.kernel < PPF_KERNEL_AUTO || d.kernel > PPF_KERNEL_FAST)
        return set_err(nullptr, PPF_EARG, "bad kernel %d", d.kernel);
    if (d.kernel == PPF_KERNEL_FAST && d.precision != PPF_F64)
        return set_err(nullptr, PPF_EARG, "the fast kernel is an FP64 form");
    if (d.kernel == PPF_KERNEL_FAST && d.params.AS != d.params.AL)
        return set_err(nullptr, PPF_EARG, "the fast FP64 form needs AS == AL");
    if (d.kernel == PPF_KERNEL_FAST && !(d.params.ML > 0.0))
        return set_err(nullptr, PPF_EARG, "the fast FP64 form needs ML > 0");
    const size_t n = (size_t)d.nz * d.ny * d.nx;
    // the step kernels address a plane with 32-bit byte offsets
    const size_t esz = d.precision == PPF_F64 ? sizeof(double) : sizeof(float);
    if ((size_t)d.ny * d.nx * esz >= ((size_t)1 << 32))
        return set_err(nullptr, PPF_EARG,
                       "an x-y plane of %d x %d sites is too large (4 GiB per plane at most)",
                       d.ny, d.nx);
    if ((n >> 36) != 0)
        return set_err(nullptr, PPF_EARG, "grid too large");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, PPF_ENODEV, "no HIP device");
    if (d.device < 0 || d.device >= ndev)
        return set_err(nullptr, PPF_ENODEV, "device %d not present (%d devices)", d.device, ndev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d.device) != hipSuccess)
        return set_err(nullptr, PPF_ENODEV, "cannot query device %d", d.device);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_err(nullptr, PPF_ENODEV, "device %d is %s, this library is built for gfx950",
                       d.device, prop.gcnArchName);
    if (hipSetDevice(d.device) != hipSuccess)
        return set_err(nullptr, PPF_EHIP, "hipSetDevice(%d) failed", d.device);

    ppf_ctx *ctx = (ppf_ctx *)calloc(1, sizeof(ppf_ctx));
    if (!ctx)
        return set_err(nullptr, PPF_ENOMEM, "host allocation failed");
    ctx->d = d;
    ctx->n = n;
    ctx->esz = d.precision == PPF_F64 ? sizeof(double) : sizeof(float);
    ctx->k = make_kconst(d.params);
    ctx->k32 = d.precision == PPF_F32 ? make_kconst32(d.params) : KConst32{};
    ctx->kf = d.kernel == PPF_KERNEL_FAST ? make_kconstc<double>(d.params) : KConstF{};
    ctx->shift_every = 50;
    ctx->eqa = (d.params.AL == d.params.AS);
    ctx->kernel = d.kernel == PPF_KERNEL_AUTO ? PPF_KERNEL_TILED : d.kernel;
    choose_tiles(ctx);
    ctx->kname = d.precision == PPF_F32                ? "k_tiled<PhysF32>"
                 : ctx->kernel == PPF_KERNEL_NAIVE ? "k_naive_f64"
                 : ctx->kernel == PPF_KERNEL_FAST  ? "k_tiled<PhysF64F>"
                                                   : "k_tiled<PhysF64>";
    int rc = PPF_OK;
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev1);
    if (e != hipSuccess) {
        set_err(nullptr, PPF_EHIP, "stream/event creation failed: %s", hipGetErrorString(e));
        ppf_destroy(ctx);
        return PPF_EHIP;
    }
    for (int c = 0; c < 2 && rc == PPF_OK; c++)
        for (int f = 0; f < 3 && rc == PPF_OK; f++) {
            e = hipMalloc(&ctx->buf[c][f], n * ctx->esz);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                rc = set_err(nullptr, PPF_ENOMEM, "hipMalloc of %zu bytes failed: %s",
                             n * ctx->esz, hipGetErrorString(e));
            }
        }
    if (rc == PPF_OK) {
        e = hipMalloc(&ctx->dres, 16 * sizeof(double));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            rc = set_err(nullptr, PPF_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
        }
    }
    if (rc != PPF_OK) {
        ppf_destroy(ctx);
        return rc;
    }
    // phasefield.c:156-167 starts every field at 1.0.
    for (int f = 0; f < 3; f++) {
        if (ctx->esz == 8)
            hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, ctx->stream,
                               (double *)ctx->buf[0][f], n, 1.0);
        else
            hipLaunchKernelGGL(k_fill<float>, dim3(1024), dim3(256), 0, ctx->stream,
                               (float *)ctx->buf[0][f], n, 1.0f);
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        set_err(nullptr, PPF_EHIP, "init fill failed: %s", hipGetErrorString(e));
        ppf_destroy(ctx);
        return PPF_EHIP;
    }
    *out = ctx;
    return PPF_OK;
}

extern "C" int ppf_destroy(ppf_ctx *ctx)
{
    if (!ctx)
        return PPF_OK;
    hipSetDevice(ctx->d.device);
    if (ctx->stream)
        hipStreamSynchronize(ctx->stream);
    for (int c = 0; c < 2; c++)
        for (int f = 0; f < 3; f++)
            if (ctx->buf[c][f])
                hipFree(ctx->buf[c][f]);
    if (ctx->partials)
        hipFree(ctx->partials);
    if (ctx->ccl_labels)
        hipFree(ctx->ccl_labels);
    if (ctx->ccl_counts)
        hipFree(ctx->ccl_counts);
    if (ctx->ccl_list)
        hipFree(ctx->ccl_list);
    if (ctx->dres)
        hipFree(ctx->dres);
    if (ctx->ev0)
        hipEventDestroy(ctx->ev0);
    if (ctx->ev1)
        hipEventDestroy(ctx->ev1);
    if (ctx->stream)
        hipStreamDestroy(ctx->stream);
    free(ctx);
    return PPF_OK;
}

// Host <-> device for one field; FP32 contexts convert on the host.
static int put_field(ppf_ctx *ctx, int f, const double *h)
{
    void *dst = ctx->buf[ctx->cur[f]][f];
    if (ctx->esz == 8) {
        HIPCHK(ctx, hipMemcpyAsync(dst, h, ctx->n * 8, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return PPF_OK;
    }
    float *tmp = (float *)malloc(ctx->n * sizeof(float));
    if (!tmp)
        return set_err(ctx, PPF_ENOMEM, "host staging allocation failed");
    for (size_t q = 0; q < ctx->n; q++)
        tmp[q] = (float)h[q];
    hipError_t e = hipMemcpyAsync(dst, tmp, ctx->n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    free(tmp);
    if (e != hipSuccess)
        return set_err(ctx, PPF_EHIP, "upload failed: %s", hipGetErrorString(e));
    return PPF_OK;
}

static int get_field(ppf_ctx *ctx, int f, double *h)
{
    const void *src = ctx->buf[ctx->cur[f]][f];
    if (ctx->esz == 8) {
        HIPCHK(ctx, hipMemcpyAsync(h, src, ctx->n * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        return PPF_OK;
    }
    float *tmp = (float *)malloc(ctx->n * sizeof(float));
    if (!tmp)
        return set_err(ctx, PPF_ENOMEM, "host staging allocation failed");
    hipError_t e = hipMemcpyAsync(tmp, src, ctx->n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
        for (size_t q = 0; q < ctx->n; q++)
            h[q] = (double)tmp[q];
    free(tmp);
    if (e != hipSuccess)
        return set_err(ctx, PPF_EHIP, "download failed: %s", hipGetErrorString(e));
    return PPF_OK;
}

extern "C" int ppf_upload(ppf_ctx *ctx, const double *s, const double *l, const double *c)
{
    if (!ctx || !s || !c)
        return ctx ? set_err(ctx, PPF_EARG, "null field pointer") : PPF_EARG;
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    int rc = put_field(ctx, 0, s);
    if (rc)
        return rc;
    if (l) {
        rc = put_field(ctx, 1, l);
    } else {
        if (ctx->esz == 8)
            hipLaunchKernelGGL(k_fill<double>, dim3(1024), dim3(256), 0, ctx->stream,
                               (double *)ctx->buf[ctx->cur[1]][1], ctx->n, 1.0);
        else
            hipLaunchKernelGGL(k_fill<float>, dim3(1024), dim3(256), 0, ctx->stream,
                               (float *)ctx->buf[ctx->cur[1]][1], ctx->n, 1.0f);
        HIPCHK(ctx, hipGetLastError());
    }
    if (rc)
        return rc;
    rc = put_field(ctx, 2, c);
    if (rc)
        return rc;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->steps = 0;
    if (ctx->d.precision == PPF_F32) {
        // the FP64 solute target is the sum of the FP64 input, not of its
        // FP32 rounding (Neumaier summation on the host)
        double sum = 0.0, comp = 0.0;
        for (size_t q = 0; q < ctx->n; q++) {
            const double t = sum + c[q];
            comp += fabs(sum) >= fabs(c[q]) ? (sum - t) + c[q] : (c[q] - t) + sum;
            sum = t;
        }
        ctx->c_target = sum + comp;
        ctx->since_shift = 0;
    }
    return PPF_OK;
}

extern "C" int ppf_download(ppf_ctx *ctx, double *s, double *l, double *c)
{
    if (!ctx)
        return PPF_EARG;
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    int rc = PPF_OK;
    if (s && (rc = get_field(ctx, 0, s)))
        return rc;
    if (l && (rc = get_field(ctx, 1, l)))
        return rc;
    if (c && (rc = get_field(ctx, 2, c)))
        return rc;
    return PPF_OK;
}

static int do_steps(ppf_ctx *ctx, int64_t nsteps, bool conly)
{
    if (!ctx)
        return PPF_EARG;
    if (nsteps < 0)
        return set_err(ctx, PPF_EARG, "negative step count");
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    const bool fused = !conly && fused_available(ctx);
    for (int64_t i = 0; i < nsteps;) {
        // a pair of steps in one launch, unless a FP32 mass shift falls
        // after the first of them (the shift must see the same state and
        // step count as with single steps)
        const bool shift_between = ctx->d.precision == PPF_F32 && ctx->shift_every > 0 &&
                                   ctx->since_shift + 1 >= ctx->shift_every;
        if (fused && nsteps - i >= 2 && !shift_between) {
            int rc = launch_step2(ctx);
            if (rc == PPF_OK)
                rc = after_step(ctx);
            if (rc == PPF_OK)
                rc = after_step(ctx);
            if (rc)
                return rc;
            i += 2;
            continue;
        }
        int rc = launch_step(ctx, conly, false);
        if (rc == PPF_OK)
            rc = after_step(ctx);
        if (rc)
            return rc;
        i++;
    }
    return PPF_OK;
}

extern "C" int ppf_step(ppf_ctx *ctx, int64_t nsteps) { return do_steps(ctx, nsteps, false); }

extern "C" int ppf_step_c_only(ppf_ctx *ctx, int64_t nsteps)
{
    return do_steps(ctx, nsteps, true);
}

extern "C" int ppf_step_report(ppf_ctx *ctx, int c_only, ppf_incsums *out)
{
    if (!ctx || !out)
        return ctx ? set_err(ctx, PPF_EARG, "null argument") : PPF_EARG;
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    const size_t nb = report_blocks(ctx);
    if (nb > (size_t)INT32_MAX)
        return set_err(ctx, PPF_EARG, "too many workgroups");
    int rc = ensure_partials(ctx, 3 * nb);
    if (rc)
        return rc;
    rc = launch_step(ctx, c_only != 0, true);
    if (rc)
        return rc;
    hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, (int)nb,
                       ctx->dres);
    HIPCHK(ctx, hipGetLastError());
    double r[3];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->dres, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    out->sum_dc2 = r[0];
    // c-only sweeps leave detaS/detaL as they were (phasefield.c:406-476):
    // report zero here; the caller keeps the earlier values if it needs them.
    out->sum_ds2 = c_only ? 0.0 : r[1];
    out->sum_dl2 = c_only ? 0.0 : r[2];
    return after_step(ctx);
}

extern "C" int ppf_reduce(ppf_ctx *ctx, ppf_stats *out)
{
    if (!ctx || !out)
        return ctx ? set_err(ctx, PPF_EARG, "null argument") : PPF_EARG;
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    const int nb = 1024;
    int rc = ensure_partials(ctx, (size_t)STAT_W * nb);
    if (rc)
        return rc;
    if (ctx->esz == 8)
        hipLaunchKernelGGL(k_stats<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double *)ctx->buf[ctx->cur[0]][0],
                           (const double *)ctx->buf[ctx->cur[1]][1],
                           (const double *)ctx->buf[ctx->cur[2]][2], ctx->n, ctx->partials);
    else
        hipLaunchKernelGGL(k_stats<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float *)ctx->buf[ctx->cur[0]][0],
                           (const float *)ctx->buf[ctx->cur[1]][1],
                           (const float *)ctx->buf[ctx->cur[2]][2], ctx->n, ctx->partials);
    HIPCHK(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, ctx->stream, ctx->partials, nb,
                       ctx->dres);
    HIPCHK(ctx, hipGetLastError());
    double r[8];
    HIPCHK(ctx, hipMemcpyAsync(r, ctx->dres, 8 * sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    out->sum_c = r[0];
    out->sum_h = r[1];
    out->n_solid = r[2];
    out->sum_s = r[3];
    out->min_c = r[4];
    out->max_c = r[5];
    out->n_nonfinite = r[6];
    out->sum_h1h = r[7];
    out->n_sites = (double)ctx->n;
    return PPF_OK;
}

extern "C" int ppf_init_spheres(ppf_ctx *ctx, int n, const double *centers_zyx,
                                const double *radii, double c_liq, double c_span)
{
    if (!ctx || n < 0 || (n > 0 && (!centers_zyx || !radii)))
        return ctx ? set_err(ctx, PPF_EARG, "bad sphere arguments") : PPF_EARG;
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    const hipError_t e =
        ctx->esz == 8
            ? init_spheres_device<double>(ctx->stream, (double *)ctx->buf[ctx->cur[0]][0],
                                          (double *)ctx->buf[ctx->cur[1]][1],
                                          (double *)ctx->buf[ctx->cur[2]][2], ctx->d.nx,
                                          ctx->d.ny, ctx->d.nz, n, centers_zyx, radii, c_liq,
                                          c_span)
            : init_spheres_device<float>(ctx->stream, (float *)ctx->buf[ctx->cur[0]][0],
                                         (float *)ctx->buf[ctx->cur[1]][1],
                                         (float *)ctx->buf[ctx->cur[2]][2], ctx->d.nx, ctx->d.ny,
                                         ctx->d.nz, n, centers_zyx, radii, c_liq, c_span);
    if (e != hipSuccess)
        (void)hipGetLastError();
    if (e != hipSuccess)
        return set_err(ctx, e == hipErrorInvalidValue ? PPF_EARG : PPF_EHIP,
                       "sphere initial condition failed: %s",
                       e == hipErrorInvalidValue ? "too many sphere-block pairs"
                                                 : hipGetErrorString(e));
    ctx->steps = 0;
    return set_target_from_device(ctx);
}

// SplitMix64 (Steele, Lea, Flood 2014) -> uniform double in [0, 1).
static inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double unit(uint64_t &x) { return (double)(splitmix64(x) >> 11) * 0x1.0p-53; }

extern "C" int ppf_spheres_generate(uint64_t seed, int n, int nz, int ny, int nx, double rmin,
                                    double rmax, double *centers_zyx, double *radii)
{
    if (n < 0 || (n > 0 && (!centers_zyx || !radii)) || nz < 1 || ny < 1 || nx < 1 ||
        !(rmin <= rmax))
        return PPF_EARG;
    uint64_t st = seed;
    for (int i = 0; i < n; i++) {
        centers_zyx[3 * i + 0] = unit(st) * nz;
        centers_zyx[3 * i + 1] = unit(st) * ny;
        centers_zyx[3 * i + 2] = unit(st) * nx;
        radii[i] = rmin + (rmax - rmin) * unit(st);
    }
    return PPF_OK;
}

extern "C" int ppf_set_tiling(ppf_ctx *ctx, int tx, int ty, int tz, int order)
{
    if (!ctx)
        return PPF_EARG;
    if (!tiled_kernel(ctx))
        return set_err(ctx, PPF_ESTATE, "tiling applies to the tiled kernels only");
    bool known = false;
    for (const TileCfg &c : k_tiles)
        known = known || (c.tx == tx && c.ty == ty);
    if (!known)
        return set_err(ctx, PPF_EARG, "tile %dx%d is not compiled", tx, ty);
    if (tz < 1)
        return set_err(ctx, PPF_EARG, "tz must be >= 1");
    if (order < 0 || order > 15)
        return set_err(ctx, PPF_EARG, "order must be 0..15");
    ctx->order = order;
    ctx->tile = {tx, ty};
    ctx->tz = tz > ctx->d.nz ? ctx->d.nz : tz;
    return PPF_OK;
}

// Time `reps` two-step launches without advancing the state; *ms per step.
static int time_fused(ppf_ctx *ctx, int reps, double *ms)
{
    int rc = PPF_OK;
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int r = 0; r < reps && rc == PPF_OK; r++)
        rc = with_policy(
            ctx, [&](auto p, const auto &k) { return launch_fused<decltype(p)>(ctx, k); });
    if (rc)
        return rc;
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = (double)f / reps / 2.0;
    return PPF_OK;
}

extern "C" int ppf_autotune(ppf_ctx *ctx, int reps, double *best_ms)
{
    if (!ctx || reps < 1 || !best_ms)
        return ctx ? set_err(ctx, PPF_EARG, "bad autotune arguments") : PPF_EARG;
    if (!tiled_kernel(ctx))
        return set_err(ctx, PPF_ESTATE, "autotuning applies to the tiled kernels only");
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    const TileCfg keep_tile = ctx->tile;
    const int keep_tz = ctx->tz, keep_order = ctx->order;
    static const int tzs[] = {16, 32, 64};
    static const int orders[] = {0, 1, 2, 4, 5, 6, 8, 10};
    double best = 1e300;
    TileCfg bt = keep_tile;
    int btz = keep_tz, bo = keep_order;
    for (const TileCfg &c : k_tiles) {
        if (c.tx > 2 * ctx->d.nx && c.tx > 16)
            continue;  // mostly idle lanes
        for (int tz : tzs)
            for (int o : orders) {
                ctx->tile = c;
                ctx->tz = tz > ctx->d.nz ? ctx->d.nz : tz;
                ctx->order = o;
                double ms = 0.0;
                int rc = time_tiling(ctx, 1, &ms);  // warm
                if (rc == PPF_OK)
                    rc = time_tiling(ctx, reps, &ms);
                if (rc) {
                    ctx->tile = keep_tile;
                    ctx->tz = keep_tz;
                    ctx->order = keep_order;
                    return rc;
                }
                if (ms < best) {
                    best = ms;
                    bt = c;
                    btz = ctx->tz;
                    bo = o;
                }
            }
    }
    ctx->tile = bt;
    ctx->tz = btz;
    ctx->order = bo;
    // the two-step kernel too, kept when faster per step
    {
        const TileCfg keep2 = ctx->tile2;
        const int keep_tz2 = ctx->tz2, keep_o2 = ctx->order2;
        static const int tzs2[] = {32, 64};
        // order bit 3 selects the ring-wave kernel, bit 2 the three-wave
        // FP64 variants (no effect in FP32)
        // bit 4 the variants with loads a plane early (fast FP64 and FP32)
        static const int orders2_64[] = {0, 2, 4, 6, 8, 10, 20, 22},
                         orders2_32[] = {0, 2, 8, 10, 16, 18};
        const bool f64 = ctx->d.precision == PPF_F64;
        const bool mirror = ctx->d.bc == PPF_BC_NOFLUX;  // bits 2 and 3 have no effect there
        const int *orders2 = f64 ? orders2_64 : orders2_32;
        const int norders2 = mirror ? 2 : f64 ? 8 : 6;
        double best2 = 1e300;
        TileCfg bt2 = keep2;
        int btz2 = keep_tz2, bo2 = keep_o2;
        // the current two-step tiling (the default, e.g. an 86-plane FP64
        // chunk outside the list below) is the first candidate
        {
            double ms = 0.0;
            int rc = time_fused(ctx, 1, &ms);
            if (rc == PPF_OK)
                rc = time_fused(ctx, reps, &ms);
            if (rc)
                return rc;
            best2 = ms;
        }
        for (const TileCfg &c : k_tiles2) {
            if (c.tx > 2 * ctx->d.nx && c.tx > 32)
                continue;
            for (int tz : tzs2)
                for (int oi = 0; oi < norders2; oi++) {
                    const int o = orders2[oi];
                    if (!fused_exists(ctx, c, o))
                        continue;  // no kernel of this kind for this tile
                    ctx->tile2 = c;
                    ctx->tz2 = tz > ctx->d.nz ? ctx->d.nz : tz;
                    ctx->order2 = o;
                    double ms = 0.0;
                    int rc = time_fused(ctx, 1, &ms);
                    if (rc == PPF_OK)
                        rc = time_fused(ctx, reps, &ms);
                    if (rc) {
                        ctx->tile2 = keep2;
                        ctx->tz2 = keep_tz2;
                        ctx->order2 = keep_o2;
                        return rc;
                    }
                    if (ms < best2) {
                        best2 = ms;
                        bt2 = c;
                        btz2 = ctx->tz2;
                        bo2 = o;
                    }
                }
        }
        ctx->tile2 = bt2;
        ctx->tz2 = btz2;
        ctx->order2 = bo2;
        ctx->fuse = best2 < best ? 2 : 1;
        best = best2 < best ? best2 : best;
    }
    *best_ms = best;
    return PPF_OK;
}

extern "C" int ppf_get_tiling(ppf_ctx *ctx, int *tx, int *ty, int *tz, int *order)
{
    if (!ctx || !tx || !ty || !tz || !order)
        return PPF_EARG;
    *order = ctx->order;
    *tx = ctx->tile.tx;
    *ty = ctx->tile.ty;
    *tz = ctx->tz;
    return PPF_OK;
}

extern "C" int ppf_set_fusion(ppf_ctx *ctx, int steps_per_launch, int tx, int ty, int tz,
                              int order)
{
    if (!ctx)
        return PPF_EARG;
    if (steps_per_launch != 1 && steps_per_launch != 2)
        return set_err(ctx, PPF_EARG, "steps per launch must be 1 or 2");
    if (steps_per_launch == 2) {
        if (!tiled_kernel(ctx))
            return set_err(ctx, PPF_ESTATE, "two steps per launch need the tiled kernels");
    }
    if (tx != 0 || ty != 0) {
        bool known = false;
        for (const TileCfg &c : k_tiles2)
            known = known || (c.tx == tx && c.ty == ty);
        if (!known)
            return set_err(ctx, PPF_EARG, "two-step tile %dx%d is not compiled", tx, ty);
    }
    if (tz < 0)
        return set_err(ctx, PPF_EARG, "tz must be >= 0");
    if (order > 31)
        return set_err(ctx, PPF_EARG, "order must be 0..31 (or negative to keep)");
    // the ring-wave kernel (order bit 3) and the mirror-boundary kernels
    // exist for some tiles only
    const TileCfg t = (tx != 0 || ty != 0) ? TileCfg{tx, ty} : ctx->tile2;
    const int o = order >= 0 ? order : ctx->order2;
    if (!fused_exists(ctx, t, o))
        return set_err(ctx, PPF_EARG, "two-step tile %dx%d has no order-%d variant%s", t.tx, t.ty,
                       o, ctx->d.bc == PPF_BC_NOFLUX ? " for no-flux boundaries" : "");
    ctx->tile2 = t;
    if (tz > 0)
        ctx->tz2 = tz > ctx->d.nz ? ctx->d.nz : tz;
    ctx->order2 = o;
    ctx->fuse = steps_per_launch;
    return PPF_OK;
}

extern "C" int ppf_get_fusion(ppf_ctx *ctx, int *steps_per_launch, int *tx, int *ty, int *tz,
                              int *order)
{
    if (!ctx || !steps_per_launch || !tx || !ty || !tz || !order)
        return PPF_EARG;
    *steps_per_launch = fused_available(ctx) ? 2 : 1;
    *tx = ctx->tile2.tx;
    *ty = ctx->tile2.ty;
    *tz = ctx->tz2;
    *order = ctx->order2;
    return PPF_OK;
}

extern "C" int ppf_particles(ppf_ctx *ctx, double threshold, ppf_particle_stats *out,
                             int64_t *sizes, int64_t cap)
{
    if (!ctx || !out || cap < 0 || (cap > 0 && !sizes))
        return ctx ? set_err(ctx, PPF_EARG, "bad particle arguments") : PPF_EARG;
    if (ctx->n >= (size_t)INT32_MAX)
        return set_err(ctx, PPF_EARG, "grid too large for 32-bit labels");
    HIPCHK(ctx, hipSetDevice(ctx->d.device));
    if (!ctx->ccl_labels) {
        HIPCHK(ctx, hipMalloc(&ctx->ccl_labels, ctx->n * sizeof(int)));
        hipError_t e = hipMalloc(&ctx->ccl_counts, ctx->n * sizeof(int));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            hipFree(ctx->ccl_labels);
            ctx->ccl_labels = nullptr;
            ctx->ccl_counts = nullptr;
            return set_err(ctx, PPF_ENOMEM, "label allocation failed: %s", hipGetErrorString(e));
        }
    }
    const unsigned g = (unsigned)((ctx->n + 255) / 256 < 65536 ? (ctx->n + 255) / 256 : 65536);
    if (ctx->esz == 8)
        hipLaunchKernelGGL(k_ccl_init<double>, dim3(g), dim3(256), 0, ctx->stream,
                           (const double *)ctx->buf[ctx->cur[0]][0], ctx->ccl_labels,
                           ctx->ccl_counts, ctx->n, threshold);
    else
        hipLaunchKernelGGL(k_ccl_init<float>, dim3(g), dim3(256), 0, ctx->stream,
                           (const float *)ctx->buf[ctx->cur[0]][0], ctx->ccl_labels,
                           ctx->ccl_counts, ctx->n, threshold);
    if (ctx->d.bc == PPF_BC_PERIODIC)
        hipLaunchKernelGGL(k_ccl_merge<0>, dim3(g), dim3(256), 0, ctx->stream, ctx->ccl_labels,
                           ctx->d.nx, ctx->d.ny, ctx->d.nz);
    else
        hipLaunchKernelGGL(k_ccl_merge<1>, dim3(g), dim3(256), 0, ctx->stream, ctx->ccl_labels,
                           ctx->d.nx, ctx->d.ny, ctx->d.nz);
    hipLaunchKernelGGL(k_ccl_flatten, dim3(g), dim3(256), 0, ctx->stream, ctx->ccl_labels,
                       ctx->ccl_counts, ctx->n);
    unsigned long long *dn = (unsigned long long *)(ctx->dres + 12);
    HIPCHK(ctx, hipMemsetAsync(dn, 0, sizeof *dn, ctx->stream));
    hipLaunchKernelGGL(k_ccl_collect, dim3(g), dim3(256), 0, ctx->stream, ctx->ccl_labels,
                       ctx->ccl_counts, ctx->n, dn, (int *)nullptr, 0ull);
    HIPCHK(ctx, hipGetLastError());
    unsigned long long nroots = 0;
    HIPCHK(ctx, hipMemcpyAsync(&nroots, dn, sizeof nroots, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<int> host(nroots);
    if (nroots > 0) {
        // the root list buffer is kept between calls and grows when needed
        if (nroots > ctx->ccl_list_cap) {
            if (ctx->ccl_list)
                HIPCHK(ctx, hipFree(ctx->ccl_list));
            ctx->ccl_list = nullptr;
            ctx->ccl_list_cap = 0;
            const size_t cap = nroots + nroots / 4 + 1024;
            HIPCHK(ctx, hipMalloc(&ctx->ccl_list, cap * sizeof(int)));
            ctx->ccl_list_cap = cap;
        }
        int *list = ctx->ccl_list;
        hipError_t e = hipMemsetAsync(dn, 0, sizeof *dn, ctx->stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_ccl_collect, dim3(g), dim3(256), 0, ctx->stream, ctx->ccl_labels,
                               ctx->ccl_counts, ctx->n, dn, list, nroots);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(host.data(), list, nroots * sizeof(int), hipMemcpyDeviceToHost,
                               ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return set_err(ctx, PPF_EHIP, "particle collection failed: %s", hipGetErrorString(e));
        }
    }
    // sizes in descending order: the order of the device list is arbitrary
    std::sort(host.begin(), host.end(), [](int a, int b) { return a > b; });
    double vol = 0.0, rsum = 0.0, rmax = 0.0;
    for (int v : host) {
        const double r = cbrt(3.0 * (double)v / (4.0 * M_PI));
        vol += v;
        rsum += r;
        rmax = r > rmax ? r : rmax;
    }
    out->n_particles = (int64_t)nroots;
    out->solid_sites = vol;
    out->mean_radius = nroots ? rsum / (double)nroots : 0.0;
    out->max_radius = rmax;
    out->mean_volume = nroots ? vol / (double)nroots : 0.0;
    for (int64_t i = 0; i < cap && i < (int64_t)nroots; i++)
        sizes[i] = host[(size_t)i];
    return PPF_OK;
}

extern "C" int ppf_set_mass_correction(ppf_ctx *ctx, int every)
{
    if (!ctx || every < 0)
        return ctx ? set_err(ctx, PPF_EARG, "bad interval %d", every) : PPF_EARG;
    if (ctx->d.precision != PPF_F32)
        return set_err(ctx, PPF_ESTATE, "mass correction applies to FP32 contexts only");
    ctx->shift_every = every;
    ctx->since_shift = 0;
    return PPF_OK;
}

extern "C" int ppf_mass_target(ppf_ctx *ctx, double *target, int64_t *n_shifts)
{
    if (!ctx || !target || !n_shifts)
        return PPF_EARG;
    if (ctx->d.precision != PPF_F32)
        return set_err(ctx, PPF_ESTATE, "mass bookkeeping applies to FP32 contexts only");
    *target = ctx->c_target;
    *n_shifts = ctx->n_shifts;
    return PPF_OK;
}

extern "C" int ppf_copy_state(ppf_ctx *dst, ppf_ctx *src)
{
    if (!dst || !src)
        return dst ? set_err(dst, PPF_EARG, "null source") : PPF_EARG;
    if (dst->n != src->n || dst->d.nz != src->d.nz || dst->d.ny != src->d.ny ||
        dst->d.device != src->d.device)
        return set_err(dst, PPF_EARG, "contexts differ in grid or device");
    HIPCHK(dst, hipSetDevice(dst->d.device));
    HIPCHK(dst, hipStreamSynchronize(src->stream));
    const unsigned g = (unsigned)((dst->n + 255) / 256 < 65536 ? (dst->n + 255) / 256 : 65536);
    for (int f = 0; f < 3; f++) {
        void *to = dst->buf[dst->cur[f]][f];
        const void *from = src->buf[src->cur[f]][f];
        if (dst->esz == src->esz)
            HIPCHK(dst, hipMemcpyAsync(to, from, dst->n * dst->esz, hipMemcpyDeviceToDevice,
                                       dst->stream));
        else if (dst->esz == 4)
            hipLaunchKernelGGL((k_convert<double, float>), dim3(g), dim3(256), 0, dst->stream,
                               (const double *)from, (float *)to, dst->n);
        else
            hipLaunchKernelGGL((k_convert<float, double>), dim3(g), dim3(256), 0, dst->stream,
                               (const float *)from, (double *)to, dst->n);
    }
    HIPCHK(dst, hipGetLastError());
    HIPCHK(dst, hipStreamSynchronize(dst->stream));
    dst->steps = src->steps;
    dst->since_shift = src->since_shift;
    if (dst->d.precision == PPF_F32 && src->d.precision == PPF_F32) {
        dst->c_target = src->c_target;
    } else if (dst->d.precision == PPF_F32) {
        // the FP64 solute target is the source's FP64 total, summed in
        // double-double on the device (not the total of its FP32 rounding)
        const int nb = 1024;
        int rc = ensure_partials(dst, (size_t)2 * nb);
        if (rc)
            return rc;
        hipLaunchKernelGGL(k_sum_dd<double>, dim3(nb), dim3(256), 0, dst->stream,
                           (const double *)src->buf[src->cur[2]][2], src->n, dst->partials);
        hipLaunchKernelGGL(k_mass_delta, dim3(1), dim3(256), 0, dst->stream, dst->partials, nb,
                           0.0, 1.0, dst->dres + 8);
        HIPCHK(dst, hipGetLastError());
        double r[2];
        HIPCHK(dst, hipMemcpyAsync(r, dst->dres + 8, sizeof r, hipMemcpyDeviceToHost,
                                   dst->stream));
        HIPCHK(dst, hipStreamSynchronize(dst->stream));
        dst->c_target = r[1];
        dst->since_shift = 0;
    }
    return PPF_OK;
}

template <typename Ta, typename Tb>
static void launch_diff(ppf_ctx *a, ppf_ctx *b, int f, double *part, int nb)
{
    hipLaunchKernelGGL((k_diff<Ta, Tb>), dim3(nb), dim3(256), 0, a->stream,
                       (const Ta *)a->buf[a->cur[f]][f], (const Tb *)b->buf[b->cur[f]][f], a->n,
                       part);
}

extern "C" int ppf_compare(ppf_ctx *a, ppf_ctx *b, ppf_diff *out)
{
    if (!a || !b || !out)
        return a ? set_err(a, PPF_EARG, "null argument") : PPF_EARG;
    if (a->n != b->n || a->d.nz != b->d.nz || a->d.ny != b->d.ny || a->d.device != b->d.device)
        return set_err(a, PPF_EARG, "contexts differ in grid or device");
    HIPCHK(a, hipSetDevice(a->d.device));
    HIPCHK(a, hipStreamSynchronize(b->stream));
    const int nb = 1024;
    int rc = ensure_partials(a, (size_t)6 * nb);
    if (rc)
        return rc;
    for (int f = 0; f < 3; f++) {
        double *part = a->partials + (size_t)2 * nb * f;
        if (a->esz == 8 && b->esz == 8)
            launch_diff<double, double>(a, b, f, part, nb);
        else if (a->esz == 8)
            launch_diff<double, float>(a, b, f, part, nb);
        else if (b->esz == 8)
            launch_diff<float, double>(a, b, f, part, nb);
        else
            launch_diff<float, float>(a, b, f, part, nb);
    }
    HIPCHK(a, hipGetLastError());
    std::vector<double> host((size_t)6 * nb);
    HIPCHK(a, hipMemcpyAsync(host.data(), a->partials, host.size() * sizeof(double),
                             hipMemcpyDeviceToHost, a->stream));
    HIPCHK(a, hipStreamSynchronize(a->stream));
    for (int f = 0; f < 3; f++) {
        double m = 0.0, t = 0.0;
        for (int j = 0; j < nb; j++) {
            const double v = host[(size_t)2 * nb * f + 2 * j];
            m = (v > m || v != v) ? v : m;
            t += host[(size_t)2 * nb * f + 2 * j + 1];
        }
        out->max_abs[f] = m;
        out->rms[f] = sqrt(t / (double)a->n);
    }
    return PPF_OK;
}

extern "C" int ppf_timer_start(ppf_ctx *ctx)
{
    if (!ctx)
        return PPF_EARG;
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return PPF_OK;
}

extern "C" int ppf_timer_stop(ppf_ctx *ctx, double *ms)
{
    if (!ctx || !ms)
        return PPF_EARG;
    HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    HIPCHK(ctx, hipEventSynchronize(ctx->ev1));
    float f = 0.0f;
    HIPCHK(ctx, hipEventElapsedTime(&f, ctx->ev0, ctx->ev1));
    *ms = (double)f;
    return PPF_OK;
}

extern "C" int ppf_sync(ppf_ctx *ctx)
{
    if (!ctx)
        return PPF_EARG;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return PPF_OK;
}

extern "C" int ppf_steps_taken(ppf_ctx *ctx, int64_t *out)
{
    if (!ctx || !out)
        return PPF_EARG;
    *out = ctx->steps;
    return PPF_OK;
}

extern "C" const char *ppf_kernel_name(ppf_ctx *ctx)
{
    if (!ctx)
        return "";
    if (fused_available(ctx))
        return ctx->d.precision == PPF_F32         ? "k_tiled2<PhysF32>"
               : ctx->kernel == PPF_KERNEL_FAST ? "k_tiled2<PhysF64F>"
                                                : "k_tiled2<PhysF64>";
    return ctx->kname;
}

extern "C" const char *ppf_last_error(ppf_ctx *ctx) { return ctx ? ctx->err : g_create_err; }

// ----------------------------------------------------------------------------
// ppf_sweep_report_noalloc: one sweep with its save-time sums on caller-owned
// device arrays, allocating nothing.  The per-workgroup partial sums and the
// three results live in `scratch`, a fixed device array of this library
// (one per GPU, held until the process ends).  Host code never passes the
// array's name to a kernel or a copy: it is a host-side stand-in, and the
// device address of each GPU's copy comes from hipGetSymbolAddress, looked
// up once per device and kept in g_scratch_ptr.
// ----------------------------------------------------------------------------
constexpr int NOALLOC_MAX_BLOCKS = 65536;  // partial-sum triples: 1.5 MiB
constexpr size_t NOALLOC_SCRATCH = 3 * (size_t)NOALLOC_MAX_BLOCKS + 8;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH];

constexpr int NOALLOC_MAX_DEV = 64;
static double *g_scratch_ptr[NOALLOC_MAX_DEV];
static std::mutex g_scratch_mu;

static hipError_t scratch_on(int device, double **ptr)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        g_scratch_ptr[device] = (double *)p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

extern "C" int ppf_sweep_report_noalloc(const double *s_in, const double *l_in,
                                        const double *c_in, double *s_out, double *l_out,
                                        double *c_out, int nz, int ny, int nx, int bc, int c_only,
                                        const ppf_params *params, int device, ppf_incsums *out)
{
    if (!s_in || !l_in || !c_in || !c_out || !params || !out || (!c_only && (!s_out || !l_out)))
        return set_err(nullptr, PPF_EARG, "null argument");
    if (nz < 1 || ny < 1 || nx < 1 || nz > MAX_DIM || ny > MAX_DIM || nx > MAX_DIM)
        return set_err(nullptr, PPF_EARG, "bad grid %d x %d x %d", nz, ny, nx);
    if ((size_t)ny * nx * sizeof(double) >= ((size_t)1 << 32))
        return set_err(nullptr, PPF_EARG,
                       "an x-y plane of %d x %d sites is too large (4 GiB per plane at most)",
                       ny, nx);
    if (bc != PPF_BC_PERIODIC && bc != PPF_BC_NOFLUX)
        return set_err(nullptr, PPF_EARG, "bad bc %d", bc);
    // the tiling a context of this grid uses for single steps (choose_tiles),
    // with longer z chunks when its workgroups would outnumber the scratch
    const TileCfg t = nx >= 64 ? TileCfg{64, 8} : nx >= 32 ? TileCfg{32, 8} : TileCfg{16, 16};
    const long ntile = (long)((nx + t.tx - 1) / t.tx) * ((ny + t.ty - 1) / t.ty);
    if (ntile > NOALLOC_MAX_BLOCKS)
        return set_err(nullptr, PPF_EARG,
                       "an x-y plane of %d x %d sites needs %ld workgroups; the scratch buffer "
                       "holds partial sums for %d",
                       ny, nx, ntile, NOALLOC_MAX_BLOCKS);
    int tz = nz < 32 ? nz : 32;
    while (ntile * ((nz + tz - 1) / tz) > NOALLOC_MAX_BLOCKS)
        tz++;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev ||
        device >= NOALLOC_MAX_DEV)
        return set_err(nullptr, PPF_ENODEV, "device %d not present", device);
    if (hipSetDevice(device) != hipSuccess)
        return set_err(nullptr, PPF_EHIP, "hipSetDevice(%d) failed", device);
    double *scr = nullptr;
    hipError_t e = scratch_on(device, &scr);
    if (e != hipSuccess)
        return set_err(nullptr, PPF_EHIP, "hipGetSymbolAddress(scratch) failed: %s",
                       hipGetErrorString(e));
    const int ntx = (nx + t.tx - 1) / t.tx, nty = (ny + t.ty - 1) / t.ty;
    const int nb = (int)(ntile * ((nz + tz - 1) / tz));
    const KConst k = make_kconst(*params);
    const bool eqa = params->AL == params->AS;
    auto launch = [&](auto p) {
        using P = decltype(p);
        tiled_fn<P> f = pick_tiled_shape<P>(t, bc, c_only != 0, true, 0);
        hipLaunchKernelGGL(f, dim3((unsigned)nb), dim3(t.tx * t.ty), 0, (hipStream_t)0, s_in, l_in,
                           c_in, s_out, l_out, c_out, nx, ny, nz, tz, ntx, nty, 0, k, scr);
    };
    if (eqa)
        launch(PhysF64<true>{});
    else
        launch(PhysF64<false>{});
    hipLaunchKernelGGL(k_sum3_final, dim3(1), dim3(256), 0, (hipStream_t)0, (const double *)scr,
                       nb, scr + 3 * (size_t)NOALLOC_MAX_BLOCKS);
    e = hipGetLastError();
    double r[3] = {0.0, 0.0, 0.0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(r, scr + 3 * (size_t)NOALLOC_MAX_BLOCKS, sizeof r,
                           hipMemcpyDeviceToHost, (hipStream_t)0);
    if (e == hipSuccess)
        e = hipStreamSynchronize((hipStream_t)0);
    if (e != hipSuccess)
        return set_err(nullptr, PPF_EHIP, "sweep failed: %s", hipGetErrorString(e));
    out->sum_dc2 = r[0];
    out->sum_ds2 = c_only ? 0.0 : r[1];
    out->sum_dl2 = c_only ? 0.0 : r[2];
    return PPF_OK;
}
