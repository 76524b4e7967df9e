// ppf_host.h -- host-side helpers shared by the translation units of
// libppf.so: the hoisted kernel constants.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ppf.h"
#include "ppf_physics.h"

// The stepping context (opaque behind include/ppf.h), shared by the
// translation units of the library.
struct TileCfg {
    int tx, ty;
};

struct ppf_ctx {
    ppf_desc d;
    KConst k;
    KConst32 k32;
    KConstF kf;  // PPF_KERNEL_FAST
    bool eqa;
    // FP32 solute bookkeeping: FP64 target of sum(c), and a mean shift of c
    // back onto it every shift_every steps (0: never).
    double c_target;
    int shift_every;
    int since_shift;
    int64_t n_shifts;
    int *ccl_labels;  // per-site labels and counts for ppf_particles, on first use
    int *ccl_counts;
    int *ccl_list;        // component roots of the last ppf_particles call
    size_t ccl_list_cap;  // in ints
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    void *buf[2][3];  // [copy][field]: etaS, etaL, c
    int cur[3];       // copy holding the current value of each field
    size_t n;
    size_t esz;
    double *partials;
    size_t partials_cap;  // in doubles
    double *dres;         // 16 doubles of results
    int64_t steps;
    int kernel;           // resolved PPF_KERNEL_*
    TileCfg tile;
    int tz;
    int order;  // workgroup order: bit 0 z-chunk fastest, bit 1 XCD-contiguous
    // Two steps per launch (tiled contexts): fuse = 2 turns it on;
    // tile2 / tz2 / order2 are that kernel's tiling.
    int fuse;
    TileCfg tile2;
    int tz2;
    int order2;
    const char *kname;
    char err[512];
};


static inline KConst make_kconst(const ppf_params &p)
{
    KConst k;
    k.AS = p.AS;
    k.AL = p.AL;
    k.dA = p.AL - p.AS;
    k.tmp5 = p.AS * p.AL;
    k.tmp6 = p.cL0 - p.cS0;
    k.cL0 = p.cL0;
    k.cS0 = p.cS0;
    k.ML = p.ML;
    k.dM = p.MS - p.ML;
    k.tmp1 = -1.0 * p.L * p.dt;
    k.tmp2 = (p.AS * p.cS0 - p.AL * p.cL0);
    k.tmp3 = p.dt * 0.5 * p.h2;
    k.hAS = 0.5 * p.AS;
    k.hAL = 0.5 * p.AL;
    k.CS = p.CS;
    k.CL = p.CL;
    k.m = p.m;
    k.g2 = 2.0 * p.gamma;
    k.kh2 = p.kappa * p.h2;
    k.K5 = k.tmp5 / p.AS;
    k.invA = 1.0 / p.AS;
    return k;
}

// The constants of the collapsed AS == AL form (ppf_physics.h, PhysC),
// computed in FP64 and rounded once to F; valid for AS == AL and ML > 0,
// which ppf_create checks for the contexts that use them.
template <typename F>
static inline KConstC<F> make_kconstc(const ppf_params &p)
{
    const double tmp1 = -1.0 * p.L * p.dt;
    const double tmp3 = p.dt * 0.5 * p.h2;
    KConstC<F> k;
    k.cL0 = (F)p.cL0;
    k.tmp6 = (F)(p.cL0 - p.cS0);
    k.rM = (F)((p.MS - p.ML) / p.ML);
    k.a1 = (F)(tmp1 * p.m);
    k.a2 = (F)(-tmp1 * p.kappa * p.h2);
    k.g2 = (F)(2.0 * p.gamma);
    k.b0 = (F)(2.0 * tmp1 * (p.CS - p.CL));
    k.b1 = (F)(2.0 * tmp1 * p.AL * (p.cL0 - p.cS0));
    k.d0 = (F)(tmp3 * p.ML * p.AL);
    return k;
}

static inline KConst32 make_kconst32(const ppf_params &p) { return make_kconstc<float>(p); }
