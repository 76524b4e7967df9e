/*
 * pf_oracle.h -- CPU oracle of the explicit Moelans phase-field step.
 * TEST INFRASTRUCTURE ONLY (see pf_oracle.c): the checker, never the product.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_BC_PERIODIC = 0, ORC_BC_NOFLUX = 1 };

/* The argument list of sweep_fields, phasefield.c:309-314, as a struct. */
typedef struct orc_params {
    double MS, ML, AS, AL, cL0, cS0, CS, CL, L, dt, h2, m, gamma, kappa;
} orc_params;

typedef struct orc_state orc_state;

void orc_params_default(orc_params *p);
orc_state *orc_new(int nz, int ny, int nx);
void orc_free(orc_state *st);
void orc_set(orc_state *st, const double *s, const double *l, const double *c);
void orc_get(const orc_state *st, double *s, double *l, double *c);
void orc_get_increments(const orc_state *st, double *ds, double *dl, double *dc);
void orc_bc_periodic(orc_state *st);
void orc_bc_noflux(orc_state *st);
void orc_sweep(orc_state *st, const orc_params *p);
void orc_sweep_c(orc_state *st, const orc_params *p);
void orc_errsums(const orc_state *st, double out[3]);
void orc_run(orc_state *st, const orc_params *p, int bc, long first_step, long nsteps,
             long n_iter_init);
int orc_run_dense(int nz, int ny, int nx, double *s, double *l, double *c, long nsteps, int bc,
                  long n_iter_init, double *sums);
int orc_run_dense_params(int nz, int ny, int nx, double *s, double *l, double *c, long nsteps,
                         int bc, long n_iter_init, const orc_params *p, double *sums);
int orc_increments_dense(int nz, int ny, int nx, const double *s, const double *l,
                         const double *c, int bc, int c_only, double *ds, double *dl,
                         double *dc);
int orc_time_dense(int nz, int ny, int nx, const double *s, const double *l, const double *c,
                   long nsteps, int bc, double *seconds);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
