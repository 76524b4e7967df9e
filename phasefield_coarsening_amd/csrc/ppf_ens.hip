// ppf_ens.hip -- batched ensembles of independent grids (C ABI in
// include/ppf.h, ppf_ens_*).
//
// All members live in one device pool, two copies of (etaS, etaL, c) each,
// and step together: one persistent launch (k_ensemble) per step walks a
// work list of (member, column tile, z chunk) items.  Each member's tile
// shape and z chunk come from the ensemble autotuner, per grid size.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <numeric>
#include <tuple>
#include <vector>

#include "ppf.h"
#include "ppf_host.h"
#include "k_ens.h"
#include "init_spheres.h"

namespace {

constexpr int ENS_NB_SUM = 64;  // workgroups per member in the solute sums

struct Tiling {
    int shape, tz;
};

const int k_shape_tx[ENS_NSHAPES] = {64, 32, 16};
const int k_shape_ty[ENS_NSHAPES] = {4, 8, 16};
const int k_shape2_tx[ENS2_NSHAPES] = {32, 16};
const int k_shape2_ty[ENS2_NSHAPES] = {8, 16};

}  // namespace

struct ppf_ens {
    int ndom, bc, prec, device;
    size_t esz;
    KConst k;
    KConst32 k32;
    bool eqa;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::vector<EnsDom> doms;
    std::vector<size_t> nsite;
    EnsDom *ddoms;
    void *pool;
    std::vector<Tiling> tiling;
    EnsItem *ditems;
    size_t items_cap;
    int nitems;
    // two steps per launch: fuse = 2 turns it on; tiling2 holds
    // the two-step shapes (ENS2_*), items2 their work list
    int fuse;
    std::vector<Tiling> tiling2;
    EnsItem *ditems2;
    size_t items2_cap;
    int nitems2;
    int *counter;
    int grid;
    int parity;
    int64_t steps;
    int shift_every, since;
    std::vector<double> targets;
    double *dtargets, *dpartials, *ddeltas, *dres;
    char err[512];
};

static thread_local char g_ens_err[512];  // last create failure, per host thread

static int eset(ppf_ens *e, int code, const char *fmt, ...)
{
    char *dst = e ? e->err : g_ens_err;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, 512, fmt, ap);
    va_end(ap);
    return code;
}

// (a failed call's error is also cleared from the thread's last error, so
// later calls do not report it)
#define EHIP(e, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            (void)hipGetLastError();                                                    \
            return eset((e), PPF_EHIP, "%s failed: %s", #call, hipGetErrorString(e_));  \
        }                                                                               \
    } while (0)

typedef void (*ens_fn64)(const EnsDom *, const EnsItem *, int, int *, int, KConst);
typedef void (*ens_fn32)(const EnsDom *, const EnsItem *, int, int *, int, KConst32);

static ens_fn64 pick64(int bc, bool eqa, bool conly)
{
#define E64(BC, E, C) \
    if (bc == BC && eqa == E && conly == C) return k_ensemble<PhysF64<E>, BC, C>;
    E64(0, true, false) E64(0, true, true) E64(0, false, false) E64(0, false, true)
    E64(1, true, false) E64(1, true, true) E64(1, false, false) E64(1, false, true)
#undef E64
    return nullptr;
}

static ens_fn32 pick32(int bc, bool conly)
{
    if (bc == 0)
        return conly ? k_ensemble<PhysF32, 0, true> : k_ensemble<PhysF32, 0, false>;
    return conly ? k_ensemble<PhysF32, 1, true> : k_ensemble<PhysF32, 1, false>;
}

// Work list for the members in `which` (all when empty), largest members
// first so the long tails start early; the z chunks of one tile are
// adjacent.
static std::vector<EnsItem> build_items(const ppf_ens *e, const std::vector<int> &which)
{
    std::vector<int> order = which;
    if (order.empty()) {
        order.resize(e->ndom);
        std::iota(order.begin(), order.end(), 0);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return e->nsite[a] > e->nsite[b]; });
    std::vector<EnsItem> items;
    for (int d : order) {
        const EnsDom &m = e->doms[d];
        const Tiling t = e->tiling[d];
        const int tx = k_shape_tx[t.shape], ty = k_shape_ty[t.shape];
        for (int y0 = 0; y0 < m.ny; y0 += ty)
            for (int x0 = 0; x0 < m.nx; x0 += tx)
                for (int z0 = 0; z0 < m.nz; z0 += t.tz)
                    items.push_back({d, t.shape, x0, y0, z0, std::min(z0 + t.tz, m.nz)});
    }
    return items;
}

static int upload_items(ppf_ens *e, const std::vector<EnsItem> &items)
{
    if (items.size() > e->items_cap) {
        if (e->ditems)
            EHIP(e, hipFree(e->ditems));
        e->ditems = nullptr;
        e->items_cap = 0;
        EHIP(e, hipMalloc(&e->ditems, items.size() * sizeof(EnsItem)));
        e->items_cap = items.size();
    }
    EHIP(e, hipMemcpyAsync(e->ditems, items.data(), items.size() * sizeof(EnsItem),
                           hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    e->nitems = (int)items.size();
    return PPF_OK;
}

// Two-step work list (tiling2 shapes), ordered as build_items.
static std::vector<EnsItem> build_items2(const ppf_ens *e, const std::vector<int> &which)
{
    std::vector<int> order = which;
    if (order.empty()) {
        order.resize(e->ndom);
        std::iota(order.begin(), order.end(), 0);
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return e->nsite[a] > e->nsite[b]; });
    std::vector<EnsItem> items;
    for (int d : order) {
        const EnsDom &m = e->doms[d];
        const Tiling t = e->tiling2[d];
        const int tx = k_shape2_tx[t.shape], ty = k_shape2_ty[t.shape];
        for (int y0 = 0; y0 < m.ny; y0 += ty)
            for (int x0 = 0; x0 < m.nx; x0 += tx)
                for (int z0 = 0; z0 < m.nz; z0 += t.tz)
                    items.push_back({d, t.shape, x0, y0, z0, std::min(z0 + t.tz, m.nz)});
    }
    return items;
}

static int upload_items2(ppf_ens *e, const std::vector<EnsItem> &items)
{
    if (items.size() > e->items2_cap) {
        if (e->ditems2)
            EHIP(e, hipFree(e->ditems2));
        e->ditems2 = nullptr;
        e->items2_cap = 0;
        EHIP(e, hipMalloc(&e->ditems2, items.size() * sizeof(EnsItem)));
        e->items2_cap = items.size();
    }
    EHIP(e, hipMemcpyAsync(e->ditems2, items.data(), items.size() * sizeof(EnsItem),
                           hipMemcpyHostToDevice, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    e->nitems2 = (int)items.size();
    return PPF_OK;
}

// One persistent two-step launch over the two-step work list: copy
// `parity` in, copy 1-parity out, two steps later.
static int launch2(ppf_ens *e)
{
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(256), 0, e->stream, e->counter, (size_t)1, 0);
    // (no-flux members: mirror ghosts in tile_march2, BC = 1)
#define ENS2(P, BC, K)                                                                       \
    hipLaunchKernelGGL((k_ensemble2<P, BC>), dim3(e->grid), dim3(256), 0, e->stream, e->ddoms, \
                       e->ditems2, e->nitems2, e->counter, e->parity, K)
    const bool nf = e->bc == PPF_BC_NOFLUX;
    if (e->prec == PPF_F32) {
        if (nf)
            ENS2(PhysF32, 1, e->k32);
        else
            ENS2(PhysF32, 0, e->k32);
    } else if (e->eqa) {
        if (nf)
            ENS2(PhysF64<true>, 1, e->k);
        else
            ENS2(PhysF64<true>, 0, e->k);
    } else {
        if (nf)
            ENS2(PhysF64<false>, 1, e->k);
        else
            ENS2(PhysF64<false>, 0, e->k);
    }
#undef ENS2
    EHIP(e, hipGetLastError());
    return PPF_OK;
}

// One persistent launch over the current work list: copy `parity` in,
// copy 1-parity out.
static int launch(ppf_ens *e, bool conly)
{
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(256), 0, e->stream, e->counter, (size_t)1, 0);
    if (e->prec == PPF_F32)
        hipLaunchKernelGGL(pick32(e->bc, conly), dim3(e->grid), dim3(256), 0, e->stream,
                           e->ddoms, e->ditems, e->nitems, e->counter, e->parity, e->k32);
    else
        hipLaunchKernelGGL(pick64(e->bc, e->eqa, conly), dim3(e->grid), dim3(256), 0, e->stream,
                           e->ddoms, e->ditems, e->nitems, e->counter, e->parity, e->k);
    EHIP(e, hipGetLastError());
    return PPF_OK;
}

static int mass_shift(ppf_ens *e)
{
    hipLaunchKernelGGL(k_ens_sum<float>, dim3(ENS_NB_SUM, e->ndom), dim3(256), 0, e->stream,
                       e->ddoms, e->parity, e->dpartials);
    hipLaunchKernelGGL(k_ens_delta, dim3(e->ndom), dim3(256), 0, e->stream, e->ddoms,
                       e->dpartials, ENS_NB_SUM, e->dtargets, e->ddeltas);
    hipLaunchKernelGGL(k_ens_shift, dim3(ENS_NB_SUM, e->ndom), dim3(256), 0, e->stream, e->ddoms,
                       e->parity, e->ddeltas, (uint64_t)e->steps);
    EHIP(e, hipGetLastError());
    return PPF_OK;
}

static bool good_dom(const ppf_ens *e, int dom) { return e && dom >= 0 && dom < e->ndom; }

extern "C" int ppf_ens_destroy(ppf_ens *e)
{
    if (!e)
        return PPF_OK;
    hipSetDevice(e->device);
    if (e->stream)
        hipStreamSynchronize(e->stream);
    void *bufs[] = {e->pool, e->ddoms, e->ditems, e->ditems2, e->counter, e->dtargets,
                    e->dpartials, e->ddeltas, e->dres};
    for (void *b : bufs)
        if (b)
            hipFree(b);
    if (e->ev0)
        hipEventDestroy(e->ev0);
    if (e->ev1)
        hipEventDestroy(e->ev1);
    if (e->stream)
        hipStreamDestroy(e->stream);
    delete e;
    return PPF_OK;
}

extern "C" int ppf_ens_create(ppf_ens **out, int ndom, const int *dims_zyx, int bc, int precision,
                              int device, const ppf_params *params)
{
    if (!out || !dims_zyx || !params || ndom < 1 || ndom > (1 << 20))
        return eset(nullptr, PPF_EARG, "bad ensemble arguments");
    *out = nullptr;
    if (bc != PPF_BC_PERIODIC && bc != PPF_BC_NOFLUX)
        return eset(nullptr, PPF_EARG, "bad bc %d", bc);
    if (precision != PPF_F64 && precision != PPF_F32)
        return eset(nullptr, PPF_EARG, "bad precision %d", precision);
    if (precision == PPF_F32 && (params->AS != params->AL || !(params->ML > 0.0)))
        return eset(nullptr, PPF_EARG, "the FP32 form needs AS == AL and ML > 0");
    for (int d = 0; d < ndom; d++)
        for (int a = 0; a < 3; a++)
            if (dims_zyx[3 * d + a] < 1 || dims_zyx[3 * d + a] > 4096)
                return eset(nullptr, PPF_EARG, "member %d has bad dims", d);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return eset(nullptr, PPF_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev)
        return eset(nullptr, PPF_ENODEV, "device %d not present", device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return eset(nullptr, PPF_ENODEV, "device %d is not gfx950", device);
    if (hipSetDevice(device) != hipSuccess)
        return eset(nullptr, PPF_EHIP, "hipSetDevice failed");

    ppf_ens *e = new ppf_ens();
    e->ndom = ndom;
    e->bc = bc;
    e->prec = precision;
    e->device = device;
    e->esz = precision == PPF_F64 ? 8 : 4;
    e->k = make_kconst(*params);
    e->k32 = precision == PPF_F32 ? make_kconst32(*params) : KConst32{};
    e->eqa = params->AS == params->AL;
    e->shift_every = 50;
    e->doms.resize(ndom);
    e->nsite.resize(ndom);
    e->tiling.resize(ndom);
    e->tiling2.resize(ndom);
    // two steps per launch by default for FP32 members (faster there, as
    // for single grids); FP64 steps one step per launch
    e->fuse = precision == PPF_F32 ? 2 : 1;
    e->targets.assign(ndom, 0.0);
    // pool: every array 256-byte aligned
    std::vector<size_t> off(ndom);
    size_t total = 0;
    for (int d = 0; d < ndom; d++) {
        EnsDom &m = e->doms[d];
        m.nz = dims_zyx[3 * d];
        m.ny = dims_zyx[3 * d + 1];
        m.nx = dims_zyx[3 * d + 2];
        e->nsite[d] = (size_t)m.nz * m.ny * m.nx;
        const size_t bytes = (e->nsite[d] * e->esz + 255) & ~(size_t)255;
        off[d] = total;
        total += 6 * bytes;
        e->tiling[d] = {m.nx >= 64 ? ENS_64x4 : (m.nx >= 32 ? ENS_32x8 : ENS_16x16),
                        m.nz < 32 ? m.nz : 32};
        e->tiling2[d] = {m.nx >= 32 ? ENS2_32x8 : ENS2_16x16, m.nz < 32 ? m.nz : 32};
    }
    hipError_t h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (h == hipSuccess)
        h = hipEventCreate(&e->ev0);
    if (h == hipSuccess)
        h = hipEventCreate(&e->ev1);
    if (h == hipSuccess)
        h = hipMalloc(&e->pool, total);
    if (h == hipSuccess)
        h = hipMalloc(&e->ddoms, ndom * sizeof(EnsDom));
    if (h == hipSuccess)
        h = hipMalloc(&e->counter, 256);
    if (h == hipSuccess)
        h = hipMalloc(&e->dtargets, ndom * sizeof(double));
    if (h == hipSuccess)
        h = hipMalloc(&e->dpartials, (size_t)2 * ENS_NB_SUM * ndom * sizeof(double));
    if (h == hipSuccess)
        h = hipMalloc(&e->ddeltas, (size_t)2 * ndom * sizeof(double));
    if (h == hipSuccess)
        h = hipMalloc(&e->dres, 16 * sizeof(double));
    if (h != hipSuccess) {
        (void)hipGetLastError();
        eset(nullptr, PPF_ENOMEM, "ensemble allocation (%zu bytes) failed: %s", total,
             hipGetErrorString(h));
        ppf_ens_destroy(e);
        return PPF_ENOMEM;
    }
    for (int d = 0; d < ndom; d++) {
        const size_t bytes = (e->nsite[d] * e->esz + 255) & ~(size_t)255;
        for (int c = 0; c < 2; c++)
            for (int f = 0; f < 3; f++)
                e->doms[d].f[c][f] = (char *)e->pool + off[d] + (size_t)(3 * c + f) * bytes;
        // phasefield.c:156-167: fields start at 1.0
        for (int f = 0; f < 3; f++) {
            if (e->esz == 8)
                hipLaunchKernelGGL(k_fill<double>, dim3(256), dim3(256), 0, e->stream,
                                   (double *)e->doms[d].f[0][f], e->nsite[d], 1.0);
            else
                hipLaunchKernelGGL(k_fill<float>, dim3(256), dim3(256), 0, e->stream,
                                   (float *)e->doms[d].f[0][f], e->nsite[d], 1.0f);
        }
    }
    h = hipMemcpyAsync(e->ddoms, e->doms.data(), ndom * sizeof(EnsDom), hipMemcpyHostToDevice,
                       e->stream);
    if (h == hipSuccess)
        h = hipStreamSynchronize(e->stream);
    // persistent grid: as many 256-thread workgroups as fit on the device
    int per_cu = 0;
    if (h == hipSuccess) {
        if (precision == PPF_F32)
            h = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu,
                                                             (const void *)pick32(bc, false), 256, 0);
        else
            h = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, (const void *)pick64(bc, e->eqa, false), 256, 0);
    }
    if (h != hipSuccess) {
        eset(nullptr, PPF_EHIP, "ensemble set-up failed: %s", hipGetErrorString(h));
        ppf_ens_destroy(e);
        return PPF_EHIP;
    }
    e->grid = prop.multiProcessorCount * (per_cu > 0 ? per_cu : 1);
    int rc = upload_items(e, build_items(e, {}));
    if (!rc)
        rc = upload_items2(e, build_items2(e, {}));
    if (rc) {
        memcpy(g_ens_err, e->err, sizeof g_ens_err);
        ppf_ens_destroy(e);
        return rc;
    }
    *out = e;
    return PPF_OK;
}

// FP64 target of one member's solute sum from its device field.
static int member_target_from_device(ppf_ens *e, int dom)
{
    if (e->prec != PPF_F32)
        return PPF_OK;
    const int nb = 256;
    double *part = nullptr;
    EHIP(e, hipMalloc(&part, 2 * nb * sizeof(double)));
    hipLaunchKernelGGL(k_sum_dd<float>, dim3(nb), dim3(256), 0, e->stream,
                       (const float *)e->doms[dom].f[e->parity][2], e->nsite[dom], part);
    hipLaunchKernelGGL(k_mass_delta, dim3(1), dim3(256), 0, e->stream, part, nb, 0.0, 1.0,
                       e->dres);
    double r[2] = {0.0, 0.0};
    hipError_t h = hipGetLastError();
    if (h == hipSuccess)
        h = hipMemcpyAsync(r, e->dres, sizeof r, hipMemcpyDeviceToHost, e->stream);
    if (h == hipSuccess)
        h = hipStreamSynchronize(e->stream);
    hipFree(part);
    if (h != hipSuccess)
        return eset(e, PPF_EHIP, "member sum failed: %s", hipGetErrorString(h));
    e->targets[dom] = r[1];
    EHIP(e, hipMemcpy(e->dtargets + dom, &r[1], sizeof(double), hipMemcpyHostToDevice));
    return PPF_OK;
}

static int put(ppf_ens *e, int dom, int f, const double *h)
{
    void *dst = e->doms[dom].f[e->parity][f];
    const size_t n = e->nsite[dom];
    if (e->esz == 8) {
        EHIP(e, hipMemcpy(dst, h, n * 8, hipMemcpyHostToDevice));
        return PPF_OK;
    }
    std::vector<float> tmp(n);
    for (size_t q = 0; q < n; q++)
        tmp[q] = (float)h[q];
    EHIP(e, hipMemcpy(dst, tmp.data(), n * 4, hipMemcpyHostToDevice));
    return PPF_OK;
}

extern "C" int ppf_ens_upload(ppf_ens *e, int dom, const double *s, const double *l,
                              const double *c)
{
    if (!good_dom(e, dom) || !s || !c)
        return e ? eset(e, PPF_EARG, "bad member or null field") : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    EHIP(e, hipStreamSynchronize(e->stream));
    int rc = put(e, dom, 0, s);
    if (!rc && l)
        rc = put(e, dom, 1, l);
    if (!rc && !l) {
        if (e->esz == 8)
            hipLaunchKernelGGL(k_fill<double>, dim3(256), dim3(256), 0, e->stream,
                               (double *)e->doms[dom].f[e->parity][1], e->nsite[dom], 1.0);
        else
            hipLaunchKernelGGL(k_fill<float>, dim3(256), dim3(256), 0, e->stream,
                               (float *)e->doms[dom].f[e->parity][1], e->nsite[dom], 1.0f);
        EHIP(e, hipStreamSynchronize(e->stream));
    }
    if (!rc)
        rc = put(e, dom, 2, c);
    if (rc || e->prec != PPF_F32)
        return rc;
    double sum = 0.0, comp = 0.0;
    for (size_t q = 0; q < e->nsite[dom]; q++) {
        const double t = sum + c[q];
        comp += fabs(sum) >= fabs(c[q]) ? (sum - t) + c[q] : (c[q] - t) + sum;
        sum = t;
    }
    e->targets[dom] = sum + comp;
    EHIP(e, hipMemcpy(e->dtargets + dom, &e->targets[dom], sizeof(double),
                      hipMemcpyHostToDevice));
    return PPF_OK;
}

template <typename H>
static int get(ppf_ens *e, int dom, int f, H *h)
{
    const void *src = e->doms[dom].f[e->parity][f];
    const size_t n = e->nsite[dom];
    if (e->esz == sizeof(H)) {
        EHIP(e, hipMemcpy(h, src, n * sizeof(H), hipMemcpyDeviceToHost));
        return PPF_OK;
    }
    if (e->esz == 8) {
        std::vector<double> tmp(n);
        EHIP(e, hipMemcpy(tmp.data(), src, n * 8, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++)
            h[q] = (H)tmp[q];
    } else {
        std::vector<float> tmp(n);
        EHIP(e, hipMemcpy(tmp.data(), src, n * 4, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; q++)
            h[q] = (H)tmp[q];
    }
    return PPF_OK;
}

template <typename H>
static int download_any(ppf_ens *e, int dom, H *s, H *l, H *c)
{
    if (!good_dom(e, dom))
        return e ? eset(e, PPF_EARG, "bad member %d", dom) : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    EHIP(e, hipStreamSynchronize(e->stream));
    int rc = PPF_OK;
    if (s && (rc = get(e, dom, 0, s)))
        return rc;
    if (l && (rc = get(e, dom, 1, l)))
        return rc;
    if (c && (rc = get(e, dom, 2, c)))
        return rc;
    return PPF_OK;
}

extern "C" int ppf_ens_download(ppf_ens *e, int dom, double *s, double *l, double *c)
{
    return download_any(e, dom, s, l, c);
}

extern "C" int ppf_ens_download_f32(ppf_ens *e, int dom, float *s, float *l, float *c)
{
    return download_any(e, dom, s, l, c);
}

extern "C" int ppf_ens_init_spheres(ppf_ens *e, int dom, int n, const double *centers_zyx,
                                    const double *radii, double c_liq, double c_span)
{
    if (!good_dom(e, dom) || n < 0 || (n > 0 && (!centers_zyx || !radii)))
        return e ? eset(e, PPF_EARG, "bad sphere arguments") : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    const EnsDom &m = e->doms[dom];
    const hipError_t h =
        e->esz == 8
            ? init_spheres_device<double>(e->stream, (double *)m.f[e->parity][0],
                                          (double *)m.f[e->parity][1],
                                          (double *)m.f[e->parity][2], m.nx, m.ny, m.nz, n,
                                          centers_zyx, radii, c_liq, c_span)
            : init_spheres_device<float>(e->stream, (float *)m.f[e->parity][0],
                                         (float *)m.f[e->parity][1], (float *)m.f[e->parity][2],
                                         m.nx, m.ny, m.nz, n, centers_zyx, radii, c_liq, c_span);
    if (h != hipSuccess)
        (void)hipGetLastError();
    if (h != hipSuccess)
        return eset(e, h == hipErrorInvalidValue ? PPF_EARG : PPF_EHIP,
                    "sphere initial condition failed: %s",
                    h == hipErrorInvalidValue ? "too many sphere-block pairs"
                                              : hipGetErrorString(h));
    return member_target_from_device(e, dom);
}

static int ens_steps(ppf_ens *e, int64_t nsteps, bool conly)
{
    if (!e || nsteps < 0)
        return e ? eset(e, PPF_EARG, "negative step count") : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    const bool fused = !conly && e->fuse == 2;
    const bool shifts = e->prec == PPF_F32 && e->shift_every > 0;
    for (int64_t i = 0; i < nsteps;) {
        // a pair in one launch unless a mass shift falls after its first step
        if (fused && nsteps - i >= 2 && !(shifts && e->since + 1 >= e->shift_every)) {
            int rc = launch2(e);
            if (rc)
                return rc;
            e->parity ^= 1;
            e->steps += 2;
            i += 2;
            if (shifts && (e->since += 2) >= e->shift_every) {
                e->since = 0;
                rc = mass_shift(e);
                if (rc)
                    return rc;
            }
            continue;
        }
        int rc = launch(e, conly);
        if (rc)
            return rc;
        e->parity ^= 1;
        e->steps++;
        i++;
        if (shifts && ++e->since >= e->shift_every) {
            e->since = 0;
            rc = mass_shift(e);
            if (rc)
                return rc;
        }
    }
    return PPF_OK;
}

extern "C" int ppf_ens_set_fusion(ppf_ens *e, int steps_per_launch)
{
    if (!e || (steps_per_launch != 1 && steps_per_launch != 2))
        return e ? eset(e, PPF_EARG, "steps per launch must be 1 or 2") : PPF_EARG;
    e->fuse = steps_per_launch;
    return PPF_OK;
}

extern "C" int ppf_ens_get_fusion(ppf_ens *e, int *steps_per_launch)
{
    if (!e || !steps_per_launch)
        return e ? eset(e, PPF_EARG, "null argument") : PPF_EARG;
    *steps_per_launch = e->fuse == 2 ? 2 : 1;
    return PPF_OK;
}

extern "C" int ppf_ens_step(ppf_ens *e, int64_t nsteps) { return ens_steps(e, nsteps, false); }

extern "C" int ppf_ens_step_c_only(ppf_ens *e, int64_t nsteps)
{
    return ens_steps(e, nsteps, true);
}

extern "C" int ppf_ens_set_tiling(ppf_ens *e, int dom, int shape, int tz)
{
    if (!good_dom(e, dom) || shape < 0 || shape >= ENS_NSHAPES || tz < 1)
        return e ? eset(e, PPF_EARG, "bad tiling") : PPF_EARG;
    e->tiling[dom] = {shape, std::min(tz, e->doms[dom].nz)};
    return upload_items(e, build_items(e, {}));
}

extern "C" int ppf_ens_set_tiling2(ppf_ens *e, int dom, int shape, int tz)
{
    if (!good_dom(e, dom) || shape < 0 || shape >= ENS2_NSHAPES || tz < 1)
        return e ? eset(e, PPF_EARG, "bad two-step tiling") : PPF_EARG;
    e->tiling2[dom] = {shape, std::min(tz, e->doms[dom].nz)};
    return upload_items2(e, build_items2(e, {}));
}

extern "C" int ppf_ens_get_tiling2(ppf_ens *e, int dom, int *shape, int *tz)
{
    if (!good_dom(e, dom) || !shape || !tz)
        return e ? eset(e, PPF_EARG, "bad arguments") : PPF_EARG;
    *shape = e->tiling2[dom].shape;
    *tz = e->tiling2[dom].tz;
    return PPF_OK;
}

extern "C" int ppf_ens_get_tiling(ppf_ens *e, int dom, int *shape, int *tz)
{
    if (!good_dom(e, dom) || !shape || !tz)
        return e ? eset(e, PPF_EARG, "bad arguments") : PPF_EARG;
    *shape = e->tiling[dom].shape;
    *tz = e->tiling[dom].tz;
    return PPF_OK;
}

// Per grid size (members with equal dims form a class): time every shape x
// tz on that class's members alone and keep the fastest.  Trial launches
// write the spare copy only, so the state does not advance.
extern "C" int ppf_ens_autotune(ppf_ens *e, int reps)
{
    if (!e || reps < 1)
        return e ? eset(e, PPF_EARG, "reps must be >= 1") : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    std::map<std::tuple<int, int, int>, std::vector<int>> classes;
    for (int d = 0; d < e->ndom; d++)
        classes[std::make_tuple(e->doms[d].nz, e->doms[d].ny, e->doms[d].nx)].push_back(d);
    static const int tzs[] = {8, 16, 32, 64};
    for (auto &kv : classes) {
        const std::vector<int> &members = kv.second;
        double best = 1e300;
        Tiling bt = e->tiling[members[0]];
        for (int shape = 0; shape < ENS_NSHAPES; shape++)
            for (int tz : tzs) {
                const Tiling t = {shape, std::min(tz, std::get<0>(kv.first))};
                for (int d : members)
                    e->tiling[d] = t;
                int rc = upload_items(e, build_items(e, members));
                if (rc)
                    return rc;
                rc = launch(e, false);  // warm
                if (rc)
                    return rc;
                EHIP(e, hipEventRecord(e->ev0, e->stream));
                for (int r = 0; r < reps && rc == PPF_OK; r++)
                    rc = launch(e, false);
                if (rc)
                    return rc;
                EHIP(e, hipEventRecord(e->ev1, e->stream));
                EHIP(e, hipEventSynchronize(e->ev1));
                float ms = 0.0f;
                EHIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
                if (ms < best) {
                    best = ms;
                    bt = t;
                }
            }
        for (int d : members)
            e->tiling[d] = bt;
    }
    int rc = upload_items(e, build_items(e, {}));
    if (rc)
        return rc;
    // the two-step shapes per class too, then the faster mode
    auto time_ms = [&](bool two, int n, float *ms) -> int {
        int r = two ? launch2(e) : launch(e, false);  // warm
        if (r)
            return r;
        EHIP(e, hipEventRecord(e->ev0, e->stream));
        for (int i = 0; i < n && r == PPF_OK; i++)
            r = two ? launch2(e) : launch(e, false);
        if (r)
            return r;
        EHIP(e, hipEventRecord(e->ev1, e->stream));
        EHIP(e, hipEventSynchronize(e->ev1));
        EHIP(e, hipEventElapsedTime(ms, e->ev0, e->ev1));
        return PPF_OK;
    };
    static const int tzs2[] = {16, 32, 64};
    for (auto &kv : classes) {
        const std::vector<int> &members = kv.second;
        double best = 1e300;
        Tiling bt = e->tiling2[members[0]];
        for (int shape = 0; shape < ENS2_NSHAPES; shape++)
            for (int tz : tzs2) {
                const Tiling t = {shape, std::min(tz, std::get<0>(kv.first))};
                for (int d : members)
                    e->tiling2[d] = t;
                if ((rc = upload_items2(e, build_items2(e, members))))
                    return rc;
                float ms = 0.0f;
                if ((rc = time_ms(true, reps, &ms)))
                    return rc;
                if (ms < best) {
                    best = ms;
                    bt = t;
                }
            }
        for (int d : members)
            e->tiling2[d] = bt;
    }
    if ((rc = upload_items2(e, build_items2(e, {}))))
        return rc;
    float ms1 = 0.0f, ms2 = 0.0f;
    if ((rc = time_ms(false, reps, &ms1)) || (rc = time_ms(true, reps, &ms2)))
        return rc;
    e->fuse = ms2 / 2.0f < ms1 ? 2 : 1;
    return PPF_OK;
}

extern "C" int ppf_ens_reduce(ppf_ens *e, int dom, ppf_stats *out)
{
    if (!good_dom(e, dom) || !out)
        return e ? eset(e, PPF_EARG, "bad arguments") : PPF_EARG;
    EHIP(e, hipSetDevice(e->device));
    const int nb = 256;
    double *part = nullptr;
    EHIP(e, hipMalloc(&part, (size_t)STAT_W * nb * sizeof(double)));
    const EnsDom &m = e->doms[dom];
    if (e->esz == 8)
        hipLaunchKernelGGL(k_stats<double>, dim3(nb), dim3(256), 0, e->stream,
                           (const double *)m.f[e->parity][0], (const double *)m.f[e->parity][1],
                           (const double *)m.f[e->parity][2], e->nsite[dom], part);
    else
        hipLaunchKernelGGL(k_stats<float>, dim3(nb), dim3(256), 0, e->stream,
                           (const float *)m.f[e->parity][0], (const float *)m.f[e->parity][1],
                           (const float *)m.f[e->parity][2], e->nsite[dom], part);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(256), 0, e->stream, part, nb, e->dres);
    double r[8];
    hipError_t h = hipGetLastError();
    if (h == hipSuccess)
        h = hipMemcpyAsync(r, e->dres, sizeof r, hipMemcpyDeviceToHost, e->stream);
    if (h == hipSuccess)
        h = hipStreamSynchronize(e->stream);
    hipFree(part);
    if (h != hipSuccess)
        return eset(e, PPF_EHIP, "reduce failed: %s", hipGetErrorString(h));
    out->sum_c = r[0];
    out->sum_h = r[1];
    out->n_solid = r[2];
    out->sum_s = r[3];
    out->min_c = r[4];
    out->max_c = r[5];
    out->n_nonfinite = r[6];
    out->sum_h1h = r[7];
    out->n_sites = (double)e->nsite[dom];
    return PPF_OK;
}

extern "C" int ppf_ens_set_mass_correction(ppf_ens *e, int every)
{
    if (!e || every < 0)
        return e ? eset(e, PPF_EARG, "bad interval") : PPF_EARG;
    if (e->prec != PPF_F32)
        return eset(e, PPF_ESTATE, "mass correction applies to FP32 ensembles only");
    e->shift_every = every;
    e->since = 0;
    return PPF_OK;
}

extern "C" int ppf_ens_mass_target(ppf_ens *e, int dom, double *target)
{
    if (!good_dom(e, dom) || !target)
        return e ? eset(e, PPF_EARG, "bad arguments") : PPF_EARG;
    *target = e->targets[dom];
    return PPF_OK;
}

extern "C" int ppf_ens_timer_start(ppf_ens *e)
{
    if (!e)
        return PPF_EARG;
    EHIP(e, hipEventRecord(e->ev0, e->stream));
    return PPF_OK;
}

extern "C" int ppf_ens_timer_stop(ppf_ens *e, double *ms)
{
    if (!e || !ms)
        return PPF_EARG;
    EHIP(e, hipEventRecord(e->ev1, e->stream));
    EHIP(e, hipEventSynchronize(e->ev1));
    float f = 0.0f;
    EHIP(e, hipEventElapsedTime(&f, e->ev0, e->ev1));
    *ms = (double)f;
    return PPF_OK;
}

extern "C" int ppf_ens_info(ppf_ens *e, int64_t *total_sites, int *nitems, int *grid)
{
    if (!e || !total_sites || !nitems || !grid)
        return PPF_EARG;
    int64_t t = 0;
    for (size_t n : e->nsite)
        t += (int64_t)n;
    *total_sites = t;
    *nitems = e->nitems;
    *grid = e->grid;
    return PPF_OK;
}

extern "C" const char *ppf_ens_last_error(ppf_ens *e) { return e ? e->err : g_ens_err; }
