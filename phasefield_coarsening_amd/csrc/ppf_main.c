/*
 * ppf_main.c -- the stepping program, a drop-in for the reference's
 * `mpirun -np N ./ppf filename` (README.md:18-22) on one MI355X.
 *
 *     ppf [options] filename
 *
 * With no options it runs the reference's shipped loop (phasefield.c:73-269):
 * no-flux boundaries (USE_BC_TYPE 1, phasefield.c:15), n_iter = 5000001,
 * n_save = 50000, n_iter_init = 1e8 (phasefield.c:104-105), stopping early
 * once err = sqrt(sum dc^2) < 1e-12 at a save (phasefield.c:99, 179).
 * Per step: ghost fill + sweep_fields, or sweep_fields_c once
 * i_iter >= n_iter_init (phasefield.c:186-218).  At i_iter > 0 with
 * i_iter % n_save == 0 or i_iter == n_iter_init (phasefield.c:221) it prints
 * the reference's two log lines to stderr and writes <stem><k>.h5, whose
 * name goes to stdout (phasefield.c:1193-1208); file k therefore holds the
 * state after i_iter + 1 sweeps.
 *
 * Options (each also readable from the environment variable in brackets):
 *   --bc=periodic|noflux      [PPF_BC]           default noflux
 *   --n-iter=N                [PPF_N_ITER]       default 5000001
 *   --n-save=N                [PPF_N_SAVE]       default 50000
 *   --n-iter-init=N           [PPF_N_ITER_INIT]  default 100000000
 *   --etal-compat             [PPF_ETAL_COMPAT=1] ignore /etaL, etaL = 1.0 as
 *                             the reference's reader leaves it
 *                             (phasefield.c:1032); default reads /etaL
 *   --device=K                [PPF_DEVICE]       default 0
 *   --stats                   [PPF_STATS=1]      one JSON line of throughput
 *                             and solute bookkeeping on stderr at the end
 *   --chunk=Z,Y,X             [PPF_CHUNK]        write the output datasets
 *                             chunked by Z x Y x X, the layout of the
 *                             reference's parallel writer (phasefield.c:
 *                             576-582); default contiguous, as its serial one
 *   --blocks=Q                [PPF_BLOCKS]       chunk = the local block of a
 *                             Q x Q x Q decomposition (Q divides each size):
 *                             the file an `mpirun -np Q^3` run writes
 *   --kernel=bitwise|fast     [PPF_KERNEL]       default bitwise: the
 *                             reference's arithmetic bit for bit; fast: FP64
 *                             with fused multiply-adds and the collapsed
 *                             AS == AL chemistry (within 1e-10 per field of
 *                             the reference after 1000 steps)
 *   --allow-nonfinite         [PPF_ALLOW_NONFINITE=1] step an input that holds
 *                             NaN or Inf values instead of refusing it
 *
 * Differences from the reference, on purpose: grid sizes come from the
 * input file (the reference compiles them in, phasefield.c:47-56); a file
 * that cannot be read is an error (the reference carries on with fields of
 * 1.0); /etaL is read from the file (the reference's readers read through
 * the closed /etaS id, so its etaL stays 1.0, phasefield.c:1026-1032;
 * --etal-compat reproduces that); an input with NaN or Inf values is
 * refused (the reference steps it); a failed output write ends the run
 * with status 1 (the reference ignores every HDF5 status,
 * phasefield.c:1052-1109); long file names do not overflow
 * (phasefield.c:107, 1201).
 */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ppf.h"
#include "ppf_io.h"

typedef struct {
    int bc;
    long n_iter, n_save, n_iter_init;
    int etal_compat;
    int device;
    int stats;
    int chunk[3];  /* output chunk dims (z, y, x); zeros: contiguous */
    int blocks;    /* > 0: chunk = dims / blocks */
    int kernel;    /* PPF_KERNEL_AUTO (bitwise) or PPF_KERNEL_FAST */
    int allow_nonfinite;
} opts_t;

static int parse_long(const char *s, long *out)
{
    char *end = NULL;
    long v = strtol(s, &end, 10);
    if (!s[0] || *end || v < 0)
        return -1;
    *out = v;
    return 0;
}

static int set_opt(opts_t *o, const char *key, const char *val)
{
    long v;
    if (!strcmp(key, "bc")) {
        if (!strcmp(val, "periodic") || !strcmp(val, "0"))
            o->bc = PPF_BC_PERIODIC;
        else if (!strcmp(val, "noflux") || !strcmp(val, "1"))
            o->bc = PPF_BC_NOFLUX;
        else
            return -1;
        return 0;
    }
    if (!strcmp(key, "n-iter"))
        return parse_long(val, &o->n_iter);
    if (!strcmp(key, "n-save")) {
        if (parse_long(val, &v) || v < 1)
            return -1;
        o->n_save = v;
        return 0;
    }
    if (!strcmp(key, "n-iter-init"))
        return parse_long(val, &o->n_iter_init);
    if (!strcmp(key, "etal-compat")) {
        o->etal_compat = strcmp(val, "0") != 0;
        return 0;
    }
    if (!strcmp(key, "device")) {
        if (parse_long(val, &v) || v > 1024)
            return -1;
        o->device = (int)v;
        return 0;
    }
    if (!strcmp(key, "stats")) {
        o->stats = strcmp(val, "0") != 0;
        return 0;
    }
    if (!strcmp(key, "chunk")) {
        int z, y, x;
        char tail;
        if (sscanf(val, "%d,%d,%d%c", &z, &y, &x, &tail) != 3 || z < 1 || y < 1 || x < 1)
            return -1;
        o->chunk[0] = z;
        o->chunk[1] = y;
        o->chunk[2] = x;
        o->blocks = 0;
        return 0;
    }
    if (!strcmp(key, "kernel")) {
        if (!strcmp(val, "bitwise"))
            o->kernel = PPF_KERNEL_AUTO;
        else if (!strcmp(val, "fast"))
            o->kernel = PPF_KERNEL_FAST;
        else
            return -1;
        return 0;
    }
    if (!strcmp(key, "allow-nonfinite")) {
        o->allow_nonfinite = strcmp(val, "0") != 0;
        return 0;
    }
    if (!strcmp(key, "blocks")) {
        if (parse_long(val, &v) || v < 1 || v > 4096)
            return -1;
        o->blocks = (int)v;
        return 0;
    }
    return -1;
}

static void env_opts(opts_t *o)
{
    static const char *map[][2] = {{"PPF_BC", "bc"},
                                   {"PPF_N_ITER", "n-iter"},
                                   {"PPF_N_SAVE", "n-save"},
                                   {"PPF_N_ITER_INIT", "n-iter-init"},
                                   {"PPF_ETAL_COMPAT", "etal-compat"},
                                   {"PPF_DEVICE", "device"},
                                   {"PPF_STATS", "stats"},
                                   {"PPF_CHUNK", "chunk"},
                                   {"PPF_BLOCKS", "blocks"},
                                   {"PPF_KERNEL", "kernel"},
                                   {"PPF_ALLOW_NONFINITE", "allow-nonfinite"}};
    for (size_t i = 0; i < sizeof map / sizeof map[0]; i++) {
        const char *v = getenv(map[i][0]);
        if (v && set_opt(o, map[i][1], v))
            fprintf(stderr, "ppf: ignoring bad %s=%s\n", map[i][0], v);
    }
}

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* phasefield.c:221 -- the save test after the sweep of iteration it. */
static int is_save(long it, const opts_t *o)
{
    return it > 0 && ((it % o->n_save) == 0 || it == o->n_iter_init);
}

static int die(const char *what, const char *msg)
{
    fprintf(stderr, "ppf: %s: %s\n", what, msg);
    return 1;
}

int main(int argc, char **argv)
{
    opts_t o = {PPF_BC_NOFLUX, 5000001L, 50000L, 100000000L, 0, 0, 0, {0, 0, 0}, 0,
                PPF_KERNEL_AUTO, 0};
    env_opts(&o);
    const char *input = NULL;
    for (int a = 1; a < argc; a++) {
        if (!strncmp(argv[a], "--", 2) && argv[a][2]) {
            char key[64];
            const char *eq = strchr(argv[a], '=');
            size_t kl = eq ? (size_t)(eq - argv[a] - 2) : strlen(argv[a] + 2);
            if (kl >= sizeof key) {
                input = NULL;
                break;
            }
            memcpy(key, argv[a] + 2, kl);
            key[kl] = 0;
            if (set_opt(&o, key, eq ? eq + 1 : "1")) {
                fprintf(stderr, "ppf: bad option %s\n", argv[a]);
                return -1;
            }
        } else if (!input) {
            input = argv[a];
        } else {
            input = NULL;
            break;
        }
    }
    /* phasefield.c:117-125 */
    char stem[4096];
    if (!input || ppfio_job_stem(input, stem, sizeof stem)) {
        fprintf(stderr, "Usage:\nppf filename\n");
        return -1;
    }
    fprintf(stderr, "The job name is %s\n", stem);

    int dims[3];
    if (ppfio_dims(input, dims))
        return die(input, ppfio_last_error());
    if (o.blocks > 0) {
        for (int i = 0; i < 3; i++) {
            if (dims[i] % o.blocks)
                return die("--blocks", "the block count must divide every grid size");
            o.chunk[i] = dims[i] / o.blocks;
        }
    }
    for (int i = 0; i < 3; i++)
        if (o.chunk[i] > dims[i])
            return die("--chunk", "a chunk dimension exceeds the grid");
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    double *s = malloc(n * sizeof(double)), *l = malloc(n * sizeof(double)),
           *c = malloc(n * sizeof(double));
    if (!s || !l || !c)
        return die("host memory", "allocation failed");
    if (ppfio_read(input, dims, s, l, c, o.etal_compat ? PPFIO_ETAL_COMPAT : 0))
        return die(input, ppfio_last_error());
    if (!o.allow_nonfinite) {
        size_t bad = 0;
        for (size_t q = 0; q < n; q++)
            bad += !isfinite(s[q]) || !isfinite(l[q]) || !isfinite(c[q]);
        if (bad) {
            fprintf(stderr, "ppf: %s: %zu of %zu sites hold NaN or Inf values "
                    "(--allow-nonfinite steps them anyway)\n", input, bad, n);
            return 1;
        }
    }

    ppf_desc d;
    memset(&d, 0, sizeof d);
    d.nz = dims[0];
    d.ny = dims[1];
    d.nx = dims[2];
    d.bc = o.bc;
    d.precision = PPF_F64;
    d.kernel = o.kernel;
    d.device = o.device;
    ppf_params_default(&d.params);
    ppf_ctx *ctx = NULL;
    if (ppf_create(&ctx, &d))
        return die("device", ppf_last_error(NULL));
    if (ppf_upload(ctx, s, l, c))
        return die("upload", ppf_last_error(ctx));
    ppf_stats st0;
    if (o.stats && ppf_reduce(ctx, &st0))
        return die("reduce", ppf_last_error(ctx));

    /* phasefield.c:178-258 */
    const double threshold = 1.0e-12;
    double error = threshold;
    double ds2 = 0.0, dl2 = 0.0; /* detaS/detaL sums of the last full sweep */
    int save_iter = 0;
    long i_iter = 0;
    double t_step = 0.0;
    while (i_iter < o.n_iter && error >= threshold) {
        const int c_only = i_iter >= o.n_iter_init;
        /* sweeps that need their increment sums: saves, and the last full
         * sweep before the c-only phase (its detaS/detaL sums are what the
         * reference keeps printing afterwards) */
        const int report = is_save(i_iter, &o) || i_iter == o.n_iter_init - 1;
        if (!report) {
            /* run up to the next sweep that needs a report, or to the end */
            long run = 1;
            while (i_iter + run < o.n_iter && !is_save(i_iter + run, &o) &&
                   i_iter + run != o.n_iter_init - 1 &&
                   (i_iter + run >= o.n_iter_init) == c_only)
                run++;
            double t0 = now_s();
            int rc = c_only ? ppf_step_c_only(ctx, run) : ppf_step(ctx, run);
            if (rc == 0)
                rc = ppf_sync(ctx);
            t_step += now_s() - t0;
            if (rc)
                return die("step", ppf_last_error(ctx));
            i_iter += run;
            continue;
        }
        ppf_incsums inc;
        double t0 = now_s();
        if (ppf_step_report(ctx, c_only, &inc))
            return die("step", ppf_last_error(ctx));
        t_step += now_s() - t0;
        if (!c_only) {
            ds2 = inc.sum_ds2;
            dl2 = inc.sum_dl2;
        }
        if (is_save(i_iter, &o)) {
            save_iter++;
            fprintf(stderr, "%ld: c=%10.4e; S=%10.4e; L=%10.4e\t", i_iter, inc.sum_dc2, ds2, dl2);
            error = sqrt(inc.sum_dc2);
            fprintf(stderr, "iter: %ld; err: %10.4e\n", i_iter, error);
            char fname[4200];
            if (ppfio_output_name(stem, save_iter, fname, sizeof fname))
                return die("output name", ppfio_last_error());
            printf("%s\n", fname);
            fflush(stdout);
            if (ppf_download(ctx, s, l, c))
                return die("download", ppf_last_error(ctx));
            if (ppfio_write_chunked(fname, dims, s, l, c, o.chunk))
                return die(fname, ppfio_last_error());
        }
        i_iter++;
    }
    if (o.stats) {
        ppf_stats st1;
        if (ppf_reduce(ctx, &st1))
            return die("reduce", ppf_last_error(ctx));
        const double glups = t_step > 0 ? (double)n * (double)i_iter / t_step / 1e9 : 0.0;
        fprintf(stderr,
                "{\"ppf_stats\": {\"grid\": [%d, %d, %d], \"steps\": %ld, \"step_seconds\": %.6f, "
                "\"glups\": %.3f, \"kernel\": \"%s\", \"sum_c0\": %.17g, \"sum_c1\": %.17g, "
                "\"solute_drift_rel\": %.3e, \"solid_fraction\": %.6f}}\n",
                dims[0], dims[1], dims[2], i_iter, t_step, glups, ppf_kernel_name(ctx), st0.sum_c,
                st1.sum_c, fabs(st1.sum_c - st0.sum_c) / fabs(st0.sum_c), st1.sum_h / st1.n_sites);
    }
    ppf_destroy(ctx);
    free(s);
    free(l);
    free(c);
    fprintf(stderr, "Total steps: %ld; error: %10.6e\n", i_iter, error);
    return 0;
}
