/*
 * pf_oracle.c -- CPU oracle for the explicit Moelans phase-field time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; nothing in the product links or calls it.
 *
 * It restates, in plain C and in the reference's own data layout (nine
 * ghost-shelled arrays, x fastest), the arithmetic of the reference program
 * phasefield.c (jijn/phasefield_coarsening):
 *
 *   orc_params_default   constants          phasefield.c:76-105
 *   orc_bc_periodic      periodic ghosts    phasefield.c:1112-1148
 *   orc_bc_noflux        mirror ghosts      phasefield.c:1151-1189
 *   orc_sweep            full sweep A/B/C   phasefield.c:301-403
 *   orc_sweep_c          c-only sweep       phasefield.c:406-476
 *   orc_errsums          save-time sums     phasefield.c:228-238
 *   orc_run              BC -> sweep loop   phasefield.c:178-218,257
 *
 * Every floating-point expression keeps the reference's operand order (C
 * evaluates left to right), and the file is compiled with -ffp-contract=off so
 * no multiply-add is fused.  With those two rules the results are bitwise
 * equal to the reference built with gcc -O2 on x86-64; the golden fixtures in
 * tests/golden pin that.
 *
 * The pass loops may run on several threads (OpenMP over the slowest axis):
 * each site is written once from values of the previous state, so the result
 * does not depend on the thread count.  The error sums stay serial because
 * their summation order is part of what they report.
 */
#include <math.h>
#include <time.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pf_oracle.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* phasefield.c:76-97 -- same literals, same expression order. */
void orc_params_default(orc_params *p)
{
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
}

struct orc_state {
    int nz, ny, nx;             /* interior sizes: NI, NJ, NK */
    size_t sj, si, n;           /* strides of the ghosted array, total count */
    double *etaS, *detaS, *etaL, *detaL, *c, *dc, *hS, *mu, *M;
};

#define AT(st, i, j, k) ((size_t)(i) * (st)->si + (size_t)(j) * (st)->sj + (size_t)(k))

orc_state *orc_new(int nz, int ny, int nx)
{
    if (nz < 1 || ny < 1 || nx < 1)
        return NULL;
    orc_state *st = (orc_state *)calloc(1, sizeof(*st));
    if (!st)
        return NULL;
    st->nz = nz;
    st->ny = ny;
    st->nx = nx;
    st->sj = (size_t)nx + 2;
    st->si = st->sj * ((size_t)ny + 2);
    st->n = st->si * ((size_t)nz + 2);
    double **arr[9] = {&st->etaS, &st->detaS, &st->etaL, &st->detaL, &st->c,
                       &st->dc, &st->hS, &st->mu, &st->M};
    for (int a = 0; a < 9; a++) {
        *arr[a] = (double *)malloc(st->n * sizeof(double));
        if (!*arr[a]) {
            orc_free(st);
            return NULL;
        }
    }
    /* phasefield.c:156-167: every array starts at 1.0 except the increments. */
    for (size_t q = 0; q < st->n; q++) {
        st->etaS[q] = 1.0;
        st->detaS[q] = 0.0;
        st->etaL[q] = 1.0;
        st->detaL[q] = 0.0;
        st->c[q] = 1.0;
        st->dc[q] = 0.0;
        st->hS[q] = 1.0;
        st->mu[q] = 1.0;
        st->M[q] = 1.0;
    }
    return st;
}

void orc_free(orc_state *st)
{
    if (!st)
        return;
    free(st->etaS);
    free(st->detaS);
    free(st->etaL);
    free(st->detaL);
    free(st->c);
    free(st->dc);
    free(st->hS);
    free(st->mu);
    free(st->M);
    free(st);
}

/* Copy interior fields in from dense [nz][ny][nx] arrays (the HDF5 corner
 * read of phasefield.c:1011-1041 lands at offset 1 in every axis).  A NULL
 * etaL leaves the array at its init value 1.0, which is what the reference's
 * reader produces (its /etaL read uses a closed dataset id,
 * phasefield.c:1032). */
void orc_set(orc_state *st, const double *s, const double *l, const double *c)
{
    size_t q = 0;
    for (int i = 1; i <= st->nz; i++)
        for (int j = 1; j <= st->ny; j++)
            for (int k = 1; k <= st->nx; k++, q++) {
                size_t a = AT(st, i, j, k);
                if (s)
                    st->etaS[a] = s[q];
                if (l)
                    st->etaL[a] = l[q];
                if (c)
                    st->c[a] = c[q];
            }
}

static void get_interior(const orc_state *st, const double *src, double *dst)
{
    size_t q = 0;
    for (int i = 1; i <= st->nz; i++)
        for (int j = 1; j <= st->ny; j++)
            for (int k = 1; k <= st->nx; k++, q++)
                dst[q] = src[AT(st, i, j, k)];
}

void orc_get(const orc_state *st, double *s, double *l, double *c)
{
    if (s)
        get_interior(st, st->etaS, s);
    if (l)
        get_interior(st, st->etaL, l);
    if (c)
        get_interior(st, st->c, c);
}

void orc_get_increments(const orc_state *st, double *ds, double *dl, double *dc)
{
    if (ds)
        get_interior(st, st->detaS, ds);
    if (dl)
        get_interior(st, st->detaL, dl);
    if (dc)
        get_interior(st, st->dc, dc);
}

/* phasefield.c:1112-1148 -- faces only; edges and corners are left alone. */
void orc_bc_periodic(orc_state *st)
{
    const int NI = st->nz, NJ = st->ny, NK = st->nx;
    double *f[3] = {st->etaS, st->etaL, st->c};
    for (int a = 0; a < 3; a++) {
        double *x = f[a];
        for (int j = 1; j <= NJ; j++)
            for (int k = 1; k <= NK; k++) {
                x[AT(st, 0, j, k)] = x[AT(st, NI, j, k)];
                x[AT(st, NI + 1, j, k)] = x[AT(st, 1, j, k)];
            }
        for (int i = 1; i <= NI; i++)
            for (int k = 1; k <= NK; k++) {
                x[AT(st, i, 0, k)] = x[AT(st, i, NJ, k)];
                x[AT(st, i, NJ + 1, k)] = x[AT(st, i, 1, k)];
            }
        for (int i = 1; i <= NI; i++)
            for (int j = 1; j <= NJ; j++) {
                x[AT(st, i, j, 0)] = x[AT(st, i, j, NK)];
                x[AT(st, i, j, NK + 1)] = x[AT(st, i, j, 1)];
            }
    }
}

/* phasefield.c:1151-1189 -- ghost face = adjacent interior face (the
 * symmetry plane lies on the cell boundary). */
void orc_bc_noflux(orc_state *st)
{
    const int NI = st->nz, NJ = st->ny, NK = st->nx;
    double *f[3] = {st->etaS, st->etaL, st->c};
    for (int a = 0; a < 3; a++) {
        double *x = f[a];
        for (int j = 1; j <= NJ; j++)
            for (int k = 1; k <= NK; k++) {
                x[AT(st, 0, j, k)] = x[AT(st, 1, j, k)];
                x[AT(st, NI + 1, j, k)] = x[AT(st, NI, j, k)];
            }
        for (int i = 1; i <= NI; i++)
            for (int k = 1; k <= NK; k++) {
                x[AT(st, i, 0, k)] = x[AT(st, i, 1, k)];
                x[AT(st, i, NJ + 1, k)] = x[AT(st, i, NJ, k)];
            }
        for (int i = 1; i <= NI; i++)
            for (int j = 1; j <= NJ; j++) {
                x[AT(st, i, j, 0)] = x[AT(st, i, j, 1)];
                x[AT(st, i, j, NK + 1)] = x[AT(st, i, j, NK)];
            }
    }
}

/* Pass A, phasefield.c:324-335: h, mu, M on every site including ghosts. */
static void pass_a(orc_state *st, const orc_params *p)
{
    const double AS = p->AS, AL = p->AL, MS = p->MS, ML = p->ML, cL0 = p->cL0;
    const double tmp5 = AS * AL;
    const double tmp6 = cL0 - p->cS0;
    const double *etaS = st->etaS, *etaL = st->etaL, *c = st->c;
    double *hS = st->hS, *mu = st->mu, *M = st->M;
    const long ni = st->nz + 2;
    const size_t si = st->si;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < ni; i++)
        for (size_t q = (size_t)i * si; q < (size_t)(i + 1) * si; q++) {
            hS[q] = etaS[q] * etaS[q] / (etaS[q] * etaS[q] + etaL[q] * etaL[q]);
            mu[q] = tmp5 / (AS + (AL - AS) * hS[q]) * (c[q] - cL0 + tmp6 * hS[q]);
            M[q] = ML + (MS - ML) * hS[q];
        }
}

/* Pass C, phasefield.c:392-401 (and :468-474 for the c-only variant). */
static void pass_c(orc_state *st, int eta_too)
{
    const long ni = st->nz + 2;
    const size_t si = st->si;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < ni; i++)
        for (size_t q = (size_t)i * si; q < (size_t)(i + 1) * si; q++) {
            if (eta_too) {
                st->etaS[q] = st->etaS[q] + st->detaS[q];
                st->etaL[q] = st->etaL[q] + st->detaL[q];
            }
            st->c[q] = st->c[q] + st->dc[q];
        }
}

/* The flux-form c increment, phasefield.c:374-384 (identical at :451-461). */
static inline double dc_site(const double *M, const double *mu, size_t q, size_t si, size_t sj,
                             double tmp3)
{
    return tmp3 * (M[q + si] * (mu[q + si] - mu[q]) -
                   M[q - si] * (mu[q] - mu[q - si]) +
                   M[q + sj] * (mu[q + sj] - mu[q]) -
                   M[q - sj] * (mu[q] - mu[q - sj]) +
                   M[q + 1] * (mu[q + 1] - mu[q]) -
                   M[q - 1] * (mu[q] - mu[q - 1]) +
                   M[q] * (mu[q + si] + mu[q - si] +
                           mu[q + sj] + mu[q - sj] +
                           mu[q + 1] + mu[q - 1] - 6.0 * mu[q]));
}

/* sweep_fields, phasefield.c:301-403. */
void orc_sweep(orc_state *st, const orc_params *p)
{
    pass_a(st, p);
    const double AS = p->AS, AL = p->AL, cS0 = p->cS0, cL0 = p->cL0;
    const double CS = p->CS, CL = p->CL, m = p->m, gamma = p->gamma;
    const double kappa = p->kappa, h2 = p->h2;
    const double tmp1 = -1.0 * p->L * p->dt;
    const double tmp2 = (AS * cS0 - AL * cL0);
    const double tmp3 = p->dt * 0.5 * h2;
    const double *etaS = st->etaS, *etaL = st->etaL, *c = st->c;
    const double *hS = st->hS, *mu = st->mu, *M = st->M;
    double *detaS = st->detaS, *detaL = st->detaL, *dc = st->dc;
    const size_t si = st->si, sj = st->sj;
    const int NI = st->nz, NJ = st->ny, NK = st->nx;
    /* Pass B, phasefield.c:347-391. */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= NI; i++)
        for (int j = 1; j <= NJ; j++)
            for (int k = 1; k <= NK; k++) {
                const size_t q = AT(st, i, j, k);
                const double tmp4 = 1.0 / (AS + (AL - AS) * hS[q]);
                const double cS = tmp4 * (AL * c[q] + (1.0 - hS[q]) * tmp2);
                const double cL = tmp4 * (AS * c[q] - hS[q] * tmp2);
                const double fS = 0.5 * AS * (cS - cS0) * (cS - cS0) + CS;
                const double fL = 0.5 * AL * (cL - cL0) * (cL - cL0) + CL;
                const double dmu = 2.0 * (etaS[q] * etaL[q]) /
                                   ((etaS[q] * etaS[q] + etaL[q] * etaL[q]) *
                                    (etaS[q] * etaS[q] + etaL[q] * etaL[q])) *
                                   (fS - fL - mu[q] * (cS - cL));
                detaS[q] = tmp1 * (m * etaS[q] * (etaS[q] * etaS[q] - 1.0 + 2.0 * gamma * etaL[q] * etaL[q])
                                   - kappa * h2 * (etaS[q + si] + etaS[q - si] +
                                                   etaS[q + sj] + etaS[q - sj] +
                                                   etaS[q + 1] + etaS[q - 1] - 6.0 * etaS[q])
                                   + etaL[q] * dmu);
                detaL[q] = tmp1 * (m * etaL[q] * (etaL[q] * etaL[q] - 1.0 + 2.0 * gamma * etaS[q] * etaS[q])
                                   - kappa * h2 * (etaL[q + si] + etaL[q - si] +
                                                   etaL[q + sj] + etaL[q - sj] +
                                                   etaL[q + 1] + etaL[q - 1] - 6.0 * etaL[q])
                                   - etaS[q] * dmu);
                dc[q] = dc_site(M, mu, q, si, sj, tmp3);
            }
    pass_c(st, 1);
}

/* sweep_fields_c, phasefield.c:406-476: eta frozen, only c moves. */
void orc_sweep_c(orc_state *st, const orc_params *p)
{
    pass_a(st, p);
    const double tmp3 = p->dt * 0.5 * p->h2;
    const size_t si = st->si, sj = st->sj;
    const int NI = st->nz, NJ = st->ny, NK = st->nx;
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= NI; i++)
        for (int j = 1; j <= NJ; j++)
            for (int k = 1; k <= NK; k++) {
                const size_t q = AT(st, i, j, k);
                st->dc[q] = dc_site(st->M, st->mu, q, si, sj, tmp3);
            }
    pass_c(st, 0);
}

/* phasefield.c:228-238: serial sums over the interior in i, j, k order. */
void orc_errsums(const orc_state *st, double out[3])
{
    double e = 0.0, e3 = 0.0, e4 = 0.0;
    for (int i = 1; i <= st->nz; i++)
        for (int j = 1; j <= st->ny; j++)
            for (int k = 1; k <= st->nx; k++) {
                const size_t q = AT(st, i, j, k);
                e += st->dc[q] * st->dc[q];
                e3 += st->detaS[q] * st->detaS[q];
                e4 += st->detaL[q] * st->detaL[q];
            }
    out[0] = e;
    out[1] = e3;
    out[2] = e4;
}

/* phasefield.c:178-218: per step, ghost fill then sweep.  Step index i uses
 * the full sweep while i < n_iter_init, the c-only sweep after
 * (phasefield.c:214-218).  first_step lets a caller continue a run. */
void orc_run(orc_state *st, const orc_params *p, int bc, long first_step, long nsteps,
             long n_iter_init)
{
    for (long it = first_step; it < first_step + nsteps; it++) {
        if (bc == ORC_BC_PERIODIC)
            orc_bc_periodic(st);
        else
            orc_bc_noflux(st);
        if (it < n_iter_init)
            orc_sweep(st, p);
        else
            orc_sweep_c(st, p);
    }
}

/* One-call form for ctypes: dense [nz][ny][nx] arrays in and out.  l may be
 * NULL (etaL stays 1.0, the reference's as-shipped reader).  sums (may be
 * NULL) receives the three save-time sums of the last step's increments.
 * Returns 0, or -1 on bad sizes / allocation failure. */
int orc_run_dense(int nz, int ny, int nx, double *s, double *l, double *c, long nsteps, int bc,
                  long n_iter_init, double *sums)
{
    orc_state *st = orc_new(nz, ny, nx);
    if (!st)
        return -1;
    orc_params p;
    orc_params_default(&p);
    orc_set(st, s, l, c);
    orc_run(st, &p, bc, 0, nsteps, n_iter_init);
    orc_get(st, s, l, c);
    if (sums)
        orc_errsums(st, sums);
    orc_free(st);
    return 0;
}

/* Same as orc_run_dense with explicit parameters. */
int orc_run_dense_params(int nz, int ny, int nx, double *s, double *l, double *c, long nsteps,
                         int bc, long n_iter_init, const orc_params *p, double *sums)
{
    orc_state *st = orc_new(nz, ny, nx);
    if (!st)
        return -1;
    orc_set(st, s, l, c);
    orc_run(st, p, bc, 0, nsteps, n_iter_init);
    orc_get(st, s, l, c);
    if (sums)
        orc_errsums(st, sums);
    orc_free(st);
    return 0;
}

/* Increments of one sweep (no state change to the caller's arrays). */
int orc_increments_dense(int nz, int ny, int nx, const double *s, const double *l,
                         const double *c, int bc, int c_only, double *ds, double *dl,
                         double *dc)
{
    orc_state *st = orc_new(nz, ny, nx);
    if (!st)
        return -1;
    orc_params p;
    orc_params_default(&p);
    orc_set(st, s, l, c);
    if (bc == ORC_BC_PERIODIC)
        orc_bc_periodic(st);
    else
        orc_bc_noflux(st);
    if (c_only)
        orc_sweep_c(st, &p);
    else
        orc_sweep(st, &p);
    orc_get_increments(st, ds, dl, dc);
    orc_free(st);
    return 0;
}

/* Timing form for the CPU baseline: like orc_run_dense, but *seconds gets
 * the wall time of the stepping loop alone (no allocation or copies). */
int orc_time_dense(int nz, int ny, int nx, const double *s, const double *l, const double *c,
                   long nsteps, int bc, double *seconds)
{
    orc_state *st = orc_new(nz, ny, nx);
    if (!st)
        return -1;
    orc_params p;
    orc_params_default(&p);
    orc_set(st, s, l, c);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    orc_run(st, &p, bc, 0, nsteps, 100000000L);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    orc_free(st);
    return 0;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
