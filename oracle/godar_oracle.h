/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for GODAR's per-timestep DEM
 * step.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it, and only as the checker.  The product library never
 * links it.
 *
 * A plain-C restatement of the reference's step (model/src/stepper.f90 and
 * the physics routines it calls) for one MPI rank and one OpenMP thread,
 * with sparse pair state instead of the reference's dense (n,n) COMMON
 * arrays.  Each function cites the reference lines it follows.
 */
#ifndef GODAR_ORACLE_H
#define GODAR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as struct godar_params in include/godar_gpu.h. */
typedef struct or_params {
    double dt, rtree, nx, ny;
    double rhoair, rhoice, rhowater;
    double Cdair, Cdwater, Csair, Cswater;
    double z0w, lat;
    double e_modul, poiss_ratio, friction_coeff, rest_coeff, sigmanc_crit;
    double eb, lambda_rb, lambda_lb, sigmatb_crit, sigmacb_crit, tau_crit;
    double bond_lim, dmax, dtd, dth;
    double ua, va, uw, vw;
    double bond_frac;
    int32_t ntree, slipping, cohesion, ridging, shelter, restart;
    uint64_t bond_seed;
} or_params;

typedef struct or_sim or_sim;

or_sim *or_create(const or_params *p, int n);
void or_destroy(or_sim *s);
void or_set_state(or_sim *s, const double *x, const double *y, const double *r,
                  const double *h, const double *theta, const double *omega,
                  const double *u, const double *v);
/* Gridded wind sampled bilinearly at floe centres every step (extension of
 * winds_currents, forcing.f90:1-29); ua/va are row-major (nyn, nxn). */
void or_set_wind_grid(or_sim *s, int nxn, int nyn, double x0, double y0,
                      double dx, double dy, const double *ua, const double *va);
/* Spatio-temporal wind (extension): nframes frames (frames, nyn, nxn) at
 * model times t0 + f frame_dt, linear in time between them; avg: averaged
 * over each floe's disc by the 7-point rule of degree 5, else at the
 * centre.  Sampled at the start of each step ((tstep - 1) dt). */
void or_set_wind_frames(or_sim *s, int nxn, int nyn, double x0, double y0, double dx,
                        double dy, int nframes, double t0, double frame_dt, int avg,
                        const double *ua, const double *va);
/* ua_i, va_i of the last step. */
void or_get_winds(or_sim *s, double *ua, double *va);
int or_step(or_sim *s, int tstep);
void or_get_state(or_sim *s, double *x, double *y, double *r, double *h,
                  double *theta, double *omega, double *u, double *v,
                  double *mass, double *inertia, double *hfa, double *hfw);
/* 33 per-floe force/diagnostic columns, row-major (33, n); column order as
 * tests/ref_harness.py FORCE_KEYS. */
void or_get_forces(or_sim *s, double *f);
void or_get_walls(or_sim *s, double *w);   /* (4, n) */
/* Candidate pairs processed in the last step, in processing order.
 * flags: bit0 contact (deltan > 0), bit1 bonded after the event,
 *        bit2 the event ridged a floe. Returns the count. */
int or_get_log(or_sim *s, int cap, int32_t *li, int32_t *lj, int32_t *lflags);
int or_get_bonds(or_sim *s, int cap, int32_t *bi, int32_t *bj);
/* Pair state for listed pairs (i<j), 18 columns row-major (18, np), order as
 * tests/ref_harness.py PAIR_KEYS. */
void or_get_pairs(or_sim *s, int np, const int32_t *pi, const int32_t *pj, double *out);
void or_get_stats(or_sim *s, int32_t *st);   /* ridge_pair, ridge_wall, ncand, ncontact */
uint32_t or_bond_threshold(double frac);
/* fit (contact_forces.f90:247-265) by itself. */
double or_fit(double xi);
int or_bond_keep(int i, int j, uint64_t seed, uint32_t thr24);

#ifdef __cplusplus
}
#endif
#endif
