/*
 * ppf.h -- C ABI of libppf.so, the MI355X (gfx950) explicit Moelans
 * phase-field stepper.
 *
 * The reference program (jijn/phasefield_coarsening, phasefield.c) steps
 * nine caller-owned, ghost-shelled static arrays on the CPU:
 *
 *   for each step:                                   phasefield.c:178-258
 *       ghost fill (periodic or no-flux)             phasefield.c:186-212
 *       sweep_fields(etaS, etaL, c, detaS, ...)      phasefield.c:215, 301-403
 *         (sweep_fields_c once i_iter >= n_iter_init phasefield.c:217, 406-476)
 *       at saves: sums of dc^2, detaS^2, detaL^2     phasefield.c:228-248
 *
 * Here the three fields live in HBM for the life of a context, the ghost
 * fill becomes index wrap (or clamp) inside the kernel, and one launch does
 * one whole sweep.  All arrays crossing this ABI are dense C order
 * [nz][ny][nx] (x fastest, the reference's k index; z is its i index,
 * phasefield.c:53-56) without ghost cells, float64, caller-owned host memory.
 *
 * Every call returns 0 or a negative PPF_E* status and never prints or
 * exits; ppf_last_error(ctx) explains the last failure on that context
 * (ppf_last_error(NULL) the calling thread's last failure of a call without
 * a context: ppf_create, ppf_sweep_report_noalloc).  One host thread
 * per context; calls are ordered on the context's own HIP stream, and the
 * calls that hand data back to the host synchronise it.
 */
#ifndef PPF_H
#define PPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PPF_OK = 0,
    PPF_EARG = -1,     /* bad argument (null pointer, size, enum value) */
    PPF_ENOMEM = -2,   /* device or host allocation failed */
    PPF_EHIP = -3,     /* a HIP runtime call failed */
    PPF_ENODEV = -4,   /* no usable gfx950 device */
    PPF_ESTATE = -5    /* call not valid in the context's state */
};

enum { PPF_BC_PERIODIC = 0, PPF_BC_NOFLUX = 1 };  /* USE_BC_TYPE, phasefield.c:15 */

/* PPF_F64 reproduces the reference bit for bit.  PPF_F32 stores and steps
 * the fields in single precision with the dimensionless form of the same
 * scheme (requires AS == AL and ML > 0), halving the bytes per site. */
enum { PPF_F64 = 0, PPF_F32 = 1 };

/* Kernel selection for PPF_F64.  PPF_KERNEL_AUTO picks the fastest exact
 * kernel; PPF_KERNEL_NAIVE is the one-thread-per-site global-memory form
 * kept as a second, independent implementation for cross-checks.
 * PPF_KERNEL_FAST (AS == AL and ML > 0 only) is FP64 arithmetic in the
 * collapsed form those constants allow: cS - cL = cS0 - cL0 and
 * fS - fL = CS - CL (phasefield.c:351-359), one reciprocal of s^2 + l^2
 * per site for h and the driving force instead of two divides, and fused
 * multiply-adds.  It is not bitwise equal to the reference; it stays
 * within 1e-10 per field of it after 1000 steps and conserves sum(c) to
 * 1e-12 relative.  Same tilings, two steps per launch and entry points as
 * PPF_KERNEL_TILED. */
enum { PPF_KERNEL_AUTO = 0, PPF_KERNEL_NAIVE = 1, PPF_KERNEL_TILED = 2, PPF_KERNEL_FAST = 3 };

/* The physical arguments of sweep_fields (phasefield.c:309-314), in that
 * order.  ppf_params_default fills the reference's values
 * (phasefield.c:76-97). */
typedef struct ppf_params {
    double MS, ML, AS, AL, cL0, cS0, CS, CL, L, dt, h2, m, gamma, kappa;
} ppf_params;

typedef struct ppf_desc {
    int nz, ny, nx;      /* interior sizes (NI, NJ, NK of phasefield.c:53-56) */
    int bc;              /* PPF_BC_* */
    int precision;       /* PPF_F64 or PPF_F32 */
    int kernel;          /* PPF_KERNEL_* */
    int device;          /* HIP device ordinal */
    int reserved;
    ppf_params params;
} ppf_desc;

/* Field statistics computed on the device (ppf_reduce). */
typedef struct ppf_stats {
    double sum_c;        /* total solute, compensated FP64 sum */
    double sum_h;        /* sum of hS = s^2/(s^2+l^2): solid phase fraction * N */
    double n_solid;      /* number of sites with etaS > 0.5 */
    double sum_s;        /* sum of etaS */
    double min_c, max_c; /* range of c */
    double n_sites;      /* nz*ny*nx */
    double n_nonfinite;  /* sites where any field is NaN or Inf */
    double sum_h1h;      /* sum of hS (1 - hS): an interface measure.  For a tanh
                          * interface phi, |grad phi| is proportional to phi (1 - phi),
                          * so 3 sum_h / sum_h1h is proportional to the mean
                          * volume-to-area radius of the solid */
} ppf_stats;

/* Sums of the last step's squared increments over the interior
 * (phasefield.c:228-238): err = sqrt(sum_dc2) is the reference's stopping
 * error (phasefield.c:247).  Summed in a fixed tree order, so they agree
 * with the reference's serial sums to rounding (~1e-15 relative). */
typedef struct ppf_incsums {
    double sum_dc2, sum_ds2, sum_dl2;
} ppf_incsums;

/* Particles of the solid phase (ppf_particles). */
typedef struct ppf_particle_stats {
    int64_t n_particles;  /* connected components of {etaS > threshold} */
    double solid_sites;   /* sites in them */
    double mean_radius;   /* mean over particles of (3V/(4 pi))^(1/3), V in sites */
    double max_radius;
    double mean_volume;
} ppf_particle_stats;

typedef struct ppf_ctx ppf_ctx;

/* Reference constants, phasefield.c:76-97. */
int ppf_params_default(ppf_params *p);

/* Allocate a context: two ping-pong copies of (etaS, etaL, c) in HBM.
 * Replaces the static arrays of phasefield.c:67-70 and their init loop
 * (phasefield.c:156-167). */
int ppf_create(ppf_ctx **out, const ppf_desc *desc);
int ppf_destroy(ppf_ctx *ctx);

/* Host -> device state.  Replaces read_hdf5's landing of the fields in the
 * arrays (phasefield.c:1011-1041).  l may be NULL: etaL is then 1.0
 * everywhere, the state the reference's reader leaves (phasefield.c:1032). */
int ppf_upload(ppf_ctx *ctx, const double *s, const double *l, const double *c);

/* Device -> host state (synchronises).  Any pointer may be NULL. */
int ppf_download(ppf_ctx *ctx, double *s, double *l, double *c);

/* nsteps sweeps, each = ghost fill + sweep_fields (phasefield.c:186-215).
 * Asynchronous on the context stream. */
int ppf_step(ppf_ctx *ctx, int64_t nsteps);

/* nsteps c-only sweeps = ghost fill + sweep_fields_c (phasefield.c:217,
 * 406-476): etaS, etaL frozen. */
int ppf_step_c_only(ppf_ctx *ctx, int64_t nsteps);

/* One step (full or c-only) that also returns the sums of its squared
 * increments, the values the reference prints at a save
 * (phasefield.c:228-248).  Synchronises. */
int ppf_step_report(ppf_ctx *ctx, int c_only, ppf_incsums *out);

/* Device reductions over the current state (synchronises). */
int ppf_reduce(ppf_ctx *ctx, ppf_stats *out);

/* Particle analysis of the current state: label the connected components
 * of {etaS > threshold} (6-neighbour connectivity, wrapping around a
 * periodic box) on the device, then report their count, total volume and
 * mean equivalent radius.  sizes (may be NULL with cap 0) receives up to cap
 * component sizes in sites, largest first.  Synchronises.  No counterpart
 * in the reference, which only reports the increment sums. */
int ppf_particles(ppf_ctx *ctx, double threshold, ppf_particle_stats *out, int64_t *sizes,
                  int64_t cap);

/* Seeded sphere initial condition on the device: etaS = max over spheres of
 * 0.5*(1 - tanh((r - R)/2)) with periodic minimum-image distance r, etaL =
 * 1 - etaS, c = c_liq + c_span*etaS.  centers_zyx holds n (z, y, x) triples
 * in cell units, radii n values.  No counterpart in the reference, which
 * takes its initial state from a MATLAB-written file (README.md:26-34). */
int ppf_init_spheres(ppf_ctx *ctx, int n, const double *centers_zyx, const double *radii,
                     double c_liq, double c_span);

/* Host-side seeded sphere list (SplitMix64): centers uniform in the box,
 * radii uniform in [rmin, rmax).  No GPU needed. */
int ppf_spheres_generate(uint64_t seed, int n, int nz, int ny, int nx, double rmin, double rmax,
                         double *centers_zyx, double *radii);

/* Column tiling of the tiled kernels: a workgroup owns a tx x ty tile of
 * the (x, y) plane and marches tz planes of z.  Compiled shapes (tx, ty):
 * (64,8) (64,16) (64,4) (32,16) (32,8) (16,16); any tz >= 1.  order picks
 * the workgroup order: bit 0 set = the z chunks of one tile are adjacent
 * (else the tiles of one z chunk are); bit 1 set = each XCD gets one
 * contiguous run of that order instead of every 8th workgroup; bit 2 set =
 * the unpipelined z march (fewer registers, loads not issued a plane
 * ahead); bit 3 set (FP64, pipelined march) = each column loaded four
 * planes ahead instead of two (more loads in flight, 12 more VGPRs).
 * Results do
 * not depend on the tiling (each site's arithmetic is fixed); only speed
 * does.  No counterpart in the reference, whose cache blocking is compiled
 * out (phasefield.c:14, 58-62). */
int ppf_set_tiling(ppf_ctx *ctx, int tx, int ty, int tz, int order);
int ppf_get_tiling(ppf_ctx *ctx, int *tx, int *ty, int *tz, int *order);

/* Two steps per launch (contexts of the tiled kernels, FP64 or FP32): one
 * z march computes step t+1 on each tile plus a one-cell ring and keeps it
 * in LDS, then step t+2 on the tile, so each site's state is read and
 * written once per two steps (24 B per site update in FP64 instead of 48).
 * Every site is updated with the same operands in the same order as by a
 * single step, so results are bit for bit those of single steps; ppf_step
 * runs pairs and a single step for an odd remainder, and never fuses across
 * an FP32 mass shift or a c-only sweep.  With no-flux (mirror) boundaries a
 * ring site or plane outside the grid is not stepped: it takes the step
 * t+1 of the site it mirrors (phasefield.c:1151-1189), written by that
 * site's thread.  On by default (FP64: 16x16 tiles at three waves per SIMD,
 * ~1.17x single steps at 512^3; FP32 periodic: 32x16 tiles with the ring in
 * waves of its own, ~1.2x).  steps_per_launch is 1 (off) or 2;
 * (tx, ty) in {(64,8), (32,8), (32,16), (64,16), (16,16)} or (0,0) to
 * keep (no-flux: (16,16) and (32,8), and (32,16) in FP32); tz > 0 or 0 to
 * keep; order bits 0-1 as ppf_set_tiling, bit 2 (FP64, tiles 32x8 and
 * 16x16) the variant that runs three waves per SIMD with no loads issued a
 * plane early, bit 3 (periodic, tiles 16x16, 32x8 and 32x16) the kernel
 * whose ring sites are updated by waves of their own after the tile's;
 * negative to keep.  PPF_EARG for a tile without the variant.
 * ppf_get_fusion reports 2 only while the context can fuse.  No counterpart
 * in the reference, which sweeps once per step (phasefield.c:178-215). */
int ppf_set_fusion(ppf_ctx *ctx, int steps_per_launch, int tx, int ty, int tz, int order);
int ppf_get_fusion(ppf_ctx *ctx, int *steps_per_launch, int *tx, int *ty, int *tz, int *order);

/* Time every compiled tiling (tile shape x tz in {16,32,64} x order in
 * {0,1,2,4,5,6,8,10}) on the current state, `reps` sweeps each, and keep the
 * fastest (ppf_get_tiling reads it back).  Also time the two-step kernel's
 * tilings (the current one, then the ppf_set_fusion shapes x tz in {32,64} x
 * order in {0,2,8,10}, and {4,6} in FP64; {0,2} with no-flux boundaries) and
 * switch two steps per launch on or off, whichever is faster per step
 * (ppf_get_fusion reads it back).  *best_ms gets the best time per step.
 * The state is not advanced: the trial sweeps write the spare copy only.
 * Synchronises. */
int ppf_autotune(ppf_ctx *ctx, int reps, double *best_ms);

/* FP32 solute bookkeeping.  An FP32 context keeps an FP64 target for the
 * total solute sum(c): the compensated sum of the FP64 input at
 * ppf_upload, or of the stored field after ppf_init_spheres.  Every `every`
 * steps (default 50; 0 switches it off) it sums c in FP64 on the device and
 * shifts c uniformly back onto the target, rounding each site
 * stochastically so that a shift far below one FP32 ulp still restores the
 * total.  ppf_mass_target reports the target and how many shifts ran.  No
 * counterpart in the reference, which is FP64 throughout and conserves
 * sum(c) by its flux form (phasefield.c:374-384). */
int ppf_set_mass_correction(ppf_ctx *ctx, int every);
int ppf_mass_target(ppf_ctx *ctx, double *target, int64_t *n_shifts);

/* Device-to-device copy of the current state (and step count, FP32
 * solute target) from src into dst; same grid and device.  Checkpoints for
 * the surrogate's exact re-run.  Contexts of different precision convert
 * (rounding to nearest, or widening); an FP32 destination of an FP64
 * source takes the FP64 total of the source's c, summed in double-double
 * on the device, as its solute target.  No counterpart in the reference. */
int ppf_copy_state(ppf_ctx *dst, ppf_ctx *src);

/* Per-field differences between the current states of two contexts of the
 * same grid and device (any precisions), computed on the device in FP64:
 * max |a - b| (NaN if any difference is NaN) and the root mean square, for
 * etaS, etaL, c in that order.  Synchronises both.  No counterpart in the
 * reference; it measures the FP32 form against FP64. */
typedef struct ppf_diff {
    double max_abs[3];
    double rms[3];
} ppf_diff;
int ppf_compare(ppf_ctx *a, ppf_ctx *b, ppf_diff *out);

/* Timing on the context stream: ppf_timer_start records a HIP event,
 * ppf_timer_stop records a second one, synchronises and returns the elapsed
 * milliseconds between them. */
int ppf_timer_start(ppf_ctx *ctx);
int ppf_timer_stop(ppf_ctx *ctx, double *ms);

/* Wait for all work queued on the context. */
int ppf_sync(ppf_ctx *ctx);

/* Steps taken since the last upload. */
int ppf_steps_taken(ppf_ctx *ctx, int64_t *out);

/* Name of the kernel ppf_step launches for multi-step runs (static string):
 * k_tiled2<...> while two steps run per launch. */
const char *ppf_kernel_name(ppf_ctx *ctx);

const char *ppf_last_error(ppf_ctx *ctx);

/* One sweep (c_only = 0: sweep_fields, 1: sweep_fields_c) of an FP64 grid
 * held in caller-owned DEVICE arrays, with the save-time sums of its squared
 * increments, that allocates nothing: no hipMalloc, hipFree or allocating
 * library routine, no stream or event creation.  No counterpart in the
 * reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).
 *   in:  s_in, l_in, c_in   device arrays [nz][ny][nx] on `device`
 *   out: s_out, l_out, c_out the new state (c_only: only c_out is written,
 *        s_out and l_out may be NULL); must not alias the inputs
 *   out: *out the sums, as ppf_step_report returns them
 * The fields equal ppf_step_report's on a context of the same grid bit for
 * bit (same kernel and tiling), and so do the sums while the grid needs at
 * most 65536 workgroups (1024^3 and below; larger grids take longer z
 * chunks and the sums differ by rounding).  The workgroup partial sums live
 * in a fixed 1.5 MiB device array of the library, one copy per GPU, held
 * until the process ends: grids whose x-y plane needs more than 65536
 * column tiles (nx*ny above ~33.5 M sites) are refused with PPF_EARG.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (two
 * host threads, or work queued on other streams) race.  Runs on the null
 * stream and synchronises it.  Errors are reported by ppf_last_error(NULL). */
int ppf_sweep_report_noalloc(const double *s_in, const double *l_in, const double *c_in,
                             double *s_out, double *l_out, double *c_out, int nz, int ny, int nx,
                             int bc, int c_only, const ppf_params *params, int device,
                             ppf_incsums *out);

/* ------------------------------------------------------------------------
 * Ensembles: many independent grids of uneven size stepped together, one
 * persistent launch per step over a work list of (member, tile, z chunk)
 * items.  Each member evolves exactly as a ppf_ctx of the same grid,
 * precision and tiling rules would (same per-site arithmetic).  No
 * counterpart in the reference, which steps one grid per MPI job.
 * ------------------------------------------------------------------------ */
typedef struct ppf_ens ppf_ens;

/* ndom members; dims_zyx holds (nz, ny, nx) per member (1..4096 each). */
int ppf_ens_create(ppf_ens **out, int ndom, const int *dims_zyx, int bc, int precision,
                   int device, const ppf_params *params);
int ppf_ens_destroy(ppf_ens *e);
int ppf_ens_upload(ppf_ens *e, int dom, const double *s, const double *l, const double *c);
int ppf_ens_download(ppf_ens *e, int dom, double *s, double *l, double *c);
/* The member's fields as float32 (snapshots for training data). */
int ppf_ens_download_f32(ppf_ens *e, int dom, float *s, float *l, float *c);
int ppf_ens_init_spheres(ppf_ens *e, int dom, int n, const double *centers_zyx,
                         const double *radii, double c_liq, double c_span);
int ppf_ens_step(ppf_ens *e, int64_t nsteps);
int ppf_ens_step_c_only(ppf_ens *e, int64_t nsteps);
/* Member tiling: shape 0 = 64x4, 1 = 32x8, 2 = 16x16 column tiles (256
 * threads each), tz planes per chunk. */
int ppf_ens_set_tiling(ppf_ens *e, int dom, int shape, int tz);
int ppf_ens_get_tiling(ppf_ens *e, int dom, int *shape, int *tz);
/* Two steps per launch (bit for bit two single steps, periodic or no-flux
 * members; pairs never straddle an FP32 mass shift): steps_per_launch 1 or
 * 2.  On by default for FP32 ensembles.  The two-step kernel has its own
 * member tiling: shape 0 = 32x8, 1 = 16x16, tz planes per chunk. */
int ppf_ens_set_fusion(ppf_ens *e, int steps_per_launch);
int ppf_ens_get_fusion(ppf_ens *e, int *steps_per_launch);
int ppf_ens_set_tiling2(ppf_ens *e, int dom, int shape, int tz);
int ppf_ens_get_tiling2(ppf_ens *e, int dom, int *shape, int *tz);
/* Per grid size, time every shape x tz in {8,16,32,64} on those members
 * alone and keep the fastest; for periodic ensembles also every two-step
 * shape x tz in {16,32,64}, then keep whichever mode (one or two steps per
 * launch) is faster per step.  The state is not advanced. */
int ppf_ens_autotune(ppf_ens *e, int reps);
int ppf_ens_reduce(ppf_ens *e, int dom, ppf_stats *out);
/* FP32 members keep an FP64 solute target each, as a ppf_ctx does. */
int ppf_ens_set_mass_correction(ppf_ens *e, int every);
int ppf_ens_mass_target(ppf_ens *e, int dom, double *target);
int ppf_ens_timer_start(ppf_ens *e);
int ppf_ens_timer_stop(ppf_ens *e, double *ms);
/* Total sites over members, work items per step, persistent workgroups. */
int ppf_ens_info(ppf_ens *e, int64_t *total_sites, int *nitems, int *grid);
const char *ppf_ens_last_error(ppf_ens *e);

#ifdef __cplusplus
}
#endif
#endif
