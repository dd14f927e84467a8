/*
 * godar_gpu.h -- C ABI of the MI355X-native GODAR step library
 * (libgodar_amd.so).
 *
 * The reference has no FFI: its driver calls one external subroutine,
 *     call stepper (tstep, restart)              model/src/godar.f90:193
 * which reads and writes the COMMON blocks of model/inc/CB_*.h.  This header
 * is what a Fortran iso_c_binding module binds instead (see
 * godar_amd/fortran/godar_gpu_mod.f90): state moves across the boundary as
 * plain arrays, device state persists between calls, and every call returns
 * an int status (0 = ok).
 *
 * Indices are 0-based; the Fortran module converts.  All arrays are host
 * pointers to n contiguous doubles unless stated otherwise.
 */
#ifndef GODAR_GPU_H
#define GODAR_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define GODAR_OK             0
#define GODAR_ERR_ARG        1   /* bad argument */
#define GODAR_ERR_HIP        2   /* HIP runtime error */
#define GODAR_ERR_CAPACITY   3   /* a per-floe list or repair table overflowed; the step was not applied */
#define GODAR_ERR_NOMEM      4   /* device allocation failed / no room for a member */
#define GODAR_ERR_STATE      5   /* call out of order (e.g. step before set_floes) */

/* Arithmetic of the force laws. */
#define GODAR_FP64           0   /* everything FP64 (the reference's precision) */
#define GODAR_MIXED          1   /* FP32 contact/bond/shelter/drag, FP64 geometry, histories,
                                    ridging bookkeeping and integration */

/* Pair kinds for godar_get_pairs. */
#define GODAR_PAIRS_CANDIDATE 0  /* pairs (i<j) processed in the last step */
#define GODAR_PAIRS_CONTACT   1  /* candidates with deltan > 0 in the last step */
#define GODAR_PAIRS_BOND      2  /* currently bonded pairs */

/*
 * Run parameters.  Mirrors the reference's parameter COMMON blocks:
 * CB_const.h (dt, rtree, ntree, densities, drag coefficients, z0w, lat,
 * disk moduli), CB_bond.h (bond constants), CB_options.h (switches),
 * CB_forcings.h (ua, va, uw, vw) and parameter.h (nx, ny).  Defaults are
 * get_default, model/src/par_get.f90:1-111.  Two workload extensions:
 * bond_frac/bond_seed (step-1 bonding of a seeded fraction of the touching
 * pairs; 1.0 = the reference rule, stepper.f90:93-101) and the wind grid.
 */
typedef struct godar_params {
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
} godar_params;

typedef struct godar_ctx godar_ctx;

/* Fill *p with the reference defaults (par_get.f90:1-111). */
int godar_params_default(godar_params *p);

/*
 * A context owns one HIP stream, device memory for up to max_floes floes in
 * up to max_members independent simulations ("members"), and the step
 * pipeline.  precision is GODAR_FP64 or GODAR_MIXED.  One host thread per
 * context; contexts are independent.
 */
int godar_create(int64_t max_floes, int32_t max_members, int32_t precision, godar_ctx **out);
int godar_destroy(godar_ctx *ctx);
const char *godar_last_error(const godar_ctx *ctx);

/*
 * Add a member (one simulation): copies the floe state to the device and
 * derives mass, inertia and freeboards (floe_properties, physics.f90:179-197)
 * as ini_get does (ini_get.f90:170-247).  tstep0 is the reference step
 * number of the member's next step (1 for a fresh run; bonds form in step 1
 * unless p->restart == 1).  *member receives the member id.  p->ntree must
 * be 1: the neighbour search uses the current positions every step, where
 * the reference with ntree > 1 searches a tree of older positions
 * (stepper.f90:44-50); other values return GODAR_ERR_ARG.
 */
int godar_member_add(godar_ctx *ctx, const godar_params *p, int32_t n,
                     const double *x, const double *y, const double *u, const double *v,
                     const double *theta, const double *omega, const double *r, const double *h,
                     int32_t tstep0, int32_t *member);
int godar_member_remove(godar_ctx *ctx, int32_t member);
/* Device bytes one member of n floes needs in this context (admission). */
int godar_member_footprint(const godar_ctx *ctx, int32_t n, int64_t *bytes);
/* Free floe slots and member slots left. */
int godar_capacity_left(const godar_ctx *ctx, int64_t *floes, int32_t *members);

/* Replace the uniform forcing (winds_currents, forcing.f90:1-29). */
int godar_set_forcing(godar_ctx *ctx, int32_t member, double ua, double va, double uw, double vw);
/* Gridded wind (extension): node (a,b) at (x0 + a*dx, y0 + b*dy), row-major
 * (nyn, nxn); sampled bilinearly at floe centres every step. */
int godar_set_wind_grid(godar_ctx *ctx, int32_t member, int32_t nxn, int32_t nyn, double x0,
                        double y0, double dx, double dy, const double *ua, const double *va);
/* Spatio-temporal gridded wind (extension; the plan of winds_currents,
 * forcing.f90:1-10): nframes frames of the grid above, frame f at model time
 * t0 + f * frame_dt (step tstep starts at model time (tstep - 1) * dt), each
 * frame's ua then va laid out as (nframes, nyn, nxn); linear in time between
 * frames, the first and last frame held outside them.  disc_average = 1: a
 * floe takes the field averaged over its disc (radius at the start of the
 * step) by the 7-point rule of degree 5 for the disc; 0: its value at the
 * centre.  godar_set_wind_grid is the case nframes = 1, disc_average = 0. */
int godar_set_wind_frames(godar_ctx *ctx, int32_t member, int32_t nxn, int32_t nyn, double x0,
                          double y0, double dx, double dy, int32_t nframes, double t0,
                          double frame_dt, int32_t disc_average, const double *ua,
                          const double *va);
/* Load bonds with their state, e.g. from a restart (i<j, 0-based).
 * state[11*k ..] = bond, damage, deltanb, deltatb, thetarelb(j,i),
 * thetarelb(i,j), rb, hb, lb, sb, ib (CB_bond.h). */
int godar_set_bonds(godar_ctx *ctx, int32_t member, int32_t nb, const int32_t *bi,
                    const int32_t *bj, const double *state);

/*
 * Checkpoints (no counterpart in the reference, whose restart reads back
 * positions, velocities, radii, thicknesses and bonds, ini_get.f90:36-39): the complete
 * state of a member -- floe state, wall histories, contact histories with the
 * last contact forces, bond records with their strain state, next step
 * number and its parameters -- as one byte block.  A member restored from it
 * continues exactly as the saved one would have (a gridded wind is not part
 * of the block: set it again).
 */
/* godar_member_restore refuses (GODAR_ERR_ARG) a block saved in a context of
 * the other precision and one whose pair records name a partner outside
 * (owner, n). */
int godar_checkpoint_size(godar_ctx *ctx, int32_t member, int64_t *bytes);
int godar_checkpoint_save(godar_ctx *ctx, int32_t member, const godar_params *p, void *buf,
                          int64_t cap);
int godar_member_restore(godar_ctx *ctx, const void *buf, int64_t size, int32_t *member);

/*
 * Advance every active member by nsteps steps: the replacement of
 * nsteps calls of stepper(tstep, restart) (stepper.f90:1-345).  A step that
 * overflows a per-floe list or the ridging-repair pools is rolled back; the
 * capacities that ran out are raised and the remaining steps run again, so
 * the result equals an uninterrupted run.  GODAR_ERR_CAPACITY is returned
 * only when a hard limit is reached (the state is then that of the last
 * complete step, see godar_member_tstep).
 *
 * godar_step_async only enqueues (no capacity growth); godar_sync waits and
 * reports GODAR_ERR_CAPACITY after rolling back to the last complete step,
 * after which godar_reserve can raise the capacities.
 */
int godar_step(godar_ctx *ctx, int32_t nsteps);
int godar_step_async(godar_ctx *ctx, int32_t nsteps);
int godar_sync(godar_ctx *ctx);
int godar_member_tstep(const godar_ctx *ctx, int32_t member, int32_t *tstep);
/* Serving (no counterpart in the reference): the member stops advancing
 * after step tlast (0: no limit); later godar_step calls leave it untouched
 * while the other members go on. */
int godar_member_run_until(godar_ctx *ctx, int32_t member, int32_t tlast);
/* Next step number of every member slot (max_members entries; -1 = free). */
int godar_member_tsteps(godar_ctx *ctx, int32_t *tsteps);
/* Per-floe list capacities (outer candidates, inner candidates, contact
 * histories, bonds per floe); 0 keeps the current value.  A capacity may
 * shrink only down to the longest list in use (else GODAR_ERR_ARG); if the
 * new lists cannot be allocated the call returns GODAR_ERR_NOMEM and the
 * context keeps its current lists. */
int godar_reserve(godar_ctx *ctx, int32_t outer, int32_t inner, int32_t hist, int32_t bonds);

/* Copy a member's floe state back: any pointer may be NULL. */
int godar_get_floes(godar_ctx *ctx, int32_t member, double *x, double *y, double *u, double *v,
                    double *theta, double *omega, double *r, double *h, double *mass,
                    double *inertia, double *hfa, double *hfw);
/* Per-floe forces of the last step, 33 columns row-major (33, n):
 * tfx tfy m fcx fcy mc fbx fby mb fax fay ma fwx fwy mw fcorx fcory
 * fx_bc fy_bc m_bc hsfa_min hsfw_min tp tsigxx tsigyy tsigxy tsigyx
 * tac pc tab pb ta_bc p_bc.
 * Requires godar_record_forces(ctx, 1) before the step; the stress
 * diagnostics tsig* (stepper.f90:238-271) are computed only in recorded
 * steps (the last contact forces they read are kept in every step). */
int godar_record_forces(godar_ctx *ctx, int32_t on);
int godar_get_forces(godar_ctx *ctx, int32_t member, double *f);
/* Wall histories (4, n): deltat_bc1 deltat_bc2 theta_bc1 theta_bc2. */
int godar_get_walls(godar_ctx *ctx, int32_t member, double *w);
/* Pair lists of a member, canonically sorted by (i, j), 0-based member-local
 * indices.  Writes up to cap pairs, *count gets the total. */
int godar_get_pairs(godar_ctx *ctx, int32_t member, int32_t kind, int32_t cap, int32_t *pi,
                    int32_t *pj, int32_t *count);
/* Bond state of the member's bonds (same order as GODAR_PAIRS_BOND),
 * 11 doubles per bond as in godar_set_bonds. */
int godar_get_bond_state(godar_ctx *ctx, int32_t member, int32_t cap, double *state,
                         int32_t *count);
/* Step statistics of the last step: [0] candidate pairs, [1] contacts,
 * [2] ridging seeds, [3] events re-evaluated by the ordered repair. */
int godar_get_stats(godar_ctx *ctx, int32_t member, int64_t *stats);
/* Totals over all members of the last step: [0] candidate pairs,
 * [1] contacts, [2] bonded pairs, [3..5] floes queued for contact, bond and
 * wall work, [6] ridging seeds and [7] events re-evaluated by the ridging
 * repair (of the last step each active member took).  `totals` holds
 * GODAR_NTOTALS entries. */
#define GODAR_NTOTALS 8
int godar_get_totals(godar_ctx *ctx, int64_t *totals);
/* godar_get_totals without any device allocation (no counterpart in the
 * reference).  For callers that cannot allocate: serving loops that poll
 * the totals between batched steps, code that must not call hipMalloc.  Its
 * working memory is a fixed 256-byte array in the library, one copy per GPU,
 * held until the process ends; every call on a GPU shares it, so calls that
 * overlap on one GPU (contexts used from several host threads) race.  Call
 * it with the context's GPU current (else GODAR_ERR_STATE).  Same result as
 * godar_get_totals. */
int godar_get_totals_noalloc(godar_ctx *ctx, int64_t *totals);

/* Pipeline tuning (no counterpart in the reference).  Threads per block
 * (64, 128 or 256) of the per-floe kernels of every stage, or of one stage
 * (stage numbers as in godar_stage_name; the sort and repair stages have
 * fixed shapes).  Defaults: 256, except 64 for the bond and wall stages and
 * 128 for the integration.  Results do not depend on block sizes. */
int godar_set_block(godar_ctx *ctx, int32_t threads);
int godar_set_stage_block(godar_ctx *ctx, int32_t stage, int32_t threads);
/* Members added afterwards whose neighbour search examines at least `floes`
 * floes per floe (3x3 cells at their grid's mean density; default 96) take
 * the dense search (supersets ordered by distance at each build, scanned by
 * a thread per floe, or two, one per half of the superset, when all dense
 * floes fit on the GPU at once) instead of the thread search (0: every
 * member, a huge value: none).  Results do not depend on it. */
int godar_set_dense_search(godar_ctx *ctx, int32_t floes);
/* Ridging repair: the number of pending events one member evaluates at once
 * (1-64, default 64; 1: one at a time) and the pending events kept in on-chip
 * memory (1-1024, default 1024) before the member continues one event at a
 * time.  Results do not depend on either. */
int godar_set_repair(godar_ctx *ctx, int32_t lanes, int32_t pending);
/* Candidate supersets of members added afterwards reach `frac` x (max r +
 * rtree) beyond each floe's query radius (0-1, default 0.1).  The cell sort
 * and search run only on steps after some floe has moved more than 0.45 of
 * that skin (or after members change); the other steps take their candidate
 * lists from the supersets.  Results do not depend on it. */
int godar_set_skin(godar_ctx *ctx, double frac);
/* Search shape of the current members and of those added afterwards: the
 * superset skin `skin_frac` x (max r + rtree) (0-1), the cells the thread
 * search scans on each side of a floe at a build (1 or 2: cells of 1/span of
 * the superset reach) and the dense-search threshold (as
 * godar_set_dense_search).  The next step builds the supersets.  Results do
 * not depend on any of them. */
int godar_set_search(godar_ctx *ctx, double skin_frac, int32_t thread_span, int32_t dense_scan);
int godar_get_search(const godar_ctx *ctx, double *skin_frac, int32_t *thread_span,
                     int32_t *dense_scan);
int godar_get_stage_blocks(const godar_ctx *ctx, int32_t *threads);   /* GODAR_NSTAGES entries */
/* Autotuner: for each tunable stage in turn, time `trial` real steps of the
 * current members with each block size and keep the fastest.  Then the
 * search shape (godar_set_search): skin 0.05, 0.1, 0.2, thread span 1, 2 and
 * dense threshold 48, 96, 192 in turn, each priced by a build step and
 * max(4, 4 trial) steps after it as t_step + (t_build - t_step) / period,
 * the period being the steps until the fastest floe (at its current speed)
 * has moved 0.45 skin, at most 5,000.  Then, unless graphs are off
 * (godar_use_graphs) or steps are being profiled, time max(32, 16 trial)
 * steps in each launch mode twice, alternating, and keep the mode with the
 * faster run.  The bond and wall stages, which run beside the contact stage,
 * are given 64-thread blocks rather than timed.  The members advance by
 * 1 + 3 * trial * (timed stages) steps, at most 6 (1 + max(4, 4 trial)) for
 * the search shape and 4 (1 + max(32, 16 trial)) for the launch mode. */
int godar_autotune(godar_ctx *ctx, int32_t trial);
/* Step launch mode (no counterpart in the reference): on (the default), each
 * step replays a captured HIP graph of the whole kernel sequence, recaptured
 * whenever members, capacities or the block size change; off, the kernels are
 * launched one by one.  Results are identical.  Steps timed by godar_profile
 * are always launched one by one. */
int godar_use_graphs(godar_ctx *ctx, int32_t on);
int godar_graphs_on(const godar_ctx *ctx, int32_t *on);   /* the current launch mode */

/*
 * Stage timing (tooling; no counterpart in the reference, which only times
 * the whole loop with omp_get_wtime, godar.f90:185-219).  When on, every
 * step records HIP events around its stages on the context's stream;
 * godar_get_profile returns the summed milliseconds and the number of timed
 * launches per stage.  godar_profile resets the totals.  Stage 6 ("unused")
 * launches nothing: the bond gather is part of the integration.
 */
#define GODAR_NSTAGES 12
int godar_profile(godar_ctx *ctx, int32_t on);
int godar_get_profile(godar_ctx *ctx, double *ms, int64_t *launches);
const char *godar_stage_name(int32_t stage);

#ifdef __cplusplus
}
#endif
#endif
