/*
 * TEST INFRASTRUCTURE ONLY -- CPU oracle for GODAR's per-timestep DEM step.
 * See godar_oracle.h.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this library, as the checker; the product never
 * links it.
 *
 * Sequential semantics restated from the reference (one rank, one thread):
 *   stepper.f90:43-343 control flow; physics.f90, contact_forces.f90,
 *   bonds.f90, ridging.f90, forcing.f90, boundaries.f90, reset.f90,
 *   integration.f90 arithmetic.  Expressions keep the reference's operand
 *   order; integer powers follow amdflang's left-to-right expansion
 *   (x**4 = ((x*x)*x)*x), and real powers x**2d0 / x**3d0 fold to the same
 *   products.  Build with -ffp-contract=off.
 *
 * Pinning (see tests/test_oracle_vs_reference.py and tests/test_oracle_laws.py):
 * every function is compared against GODAR's own routines compiled from the
 * reference sources (oracle/_ref) on trajectories with contacts, bonds,
 * walls, ridging and sheltering active, plus closed-form checks.  The
 * neighbour predicate replaces the external coretran k-d tree; its radius
 * semantics (dist <= r(i)+rtree, ascending distance, ties by index) are this
 * project's definition ("parity unpinned" against coretran itself, which is
 * not available).
 */
#define _GNU_SOURCE
#include "godar_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* sparse pair state (replaces the dense (n,n) COMMON arrays)          */
/* ------------------------------------------------------------------ */
typedef struct pstate {
    uint64_t key;                 /* (i << 32) | j, 0-based, i < j; 0 = empty slot marker via used */
    int used;
    /* contact history (CB_variables.h: deltat, thetarelc) */
    double deltat, thetarelc;
    /* stale-by-design per-pair forces used by the stress diagnostics */
    double fcn, fct, fcr, mcc;
    /* bond state (CB_bond.h) */
    int bond;
    double damage, deltanb, deltatb, trb_ji, trb_ij, rb, hb, lb, sb, ib;
    double fbn, fbt, mbb_ji, mbb_ij;
} pstate;

typedef struct cand { double d; int j; } cand;

struct or_sim {
    or_params p;
    int n;
    double pi, ec, gc, beta, sinlat, log10z0;
    /* per floe (CB_variables.h:5-17) */
    double *x, *y, *r, *h, *hfa, *hfw, *mass, *inertia, *u, *v, *theta, *omega;
    /* accumulators (CB_variables.h, CB_bond.h, CB_forcings.h, CB_diagnostics.h) */
    double *fcx, *fcy, *mc, *fbx, *fby, *mb;
    double *fax, *fay, *fwx, *fwy, *ma, *mw, *fcorx, *fcory;
    double *fn_bc, *ft_bc, *fr_bc, *fx_bc, *fy_bc, *mc_bc, *m_bc;
    double *deltat_bc1, *deltat_bc2, *theta_bc1, *theta_bc2, *delt_ridge_bc;
    double *tfx, *tfy, *m;
    double *loc_a, *loc_w, *thr_a, *thr_w;     /* local_hsf*_min and the per-thread minima */
    double *ua_i, *va_i, *uw_i, *vw_i;
    double *sigxx, *sigyy, *sigxy, *sigyx;
    double *sigxx_bc, *sigyy_bc, *sigxy_bc, *sigyx_bc;
    double *sigxx_aw, *sigyy_aw, *sigxy_aw, *sigyx_aw;
    double *tac, *tab, *pc, *pb, *tp, *ta_bc, *p_bc;
    double *tsig[4];
    /* wind grid */
    int wg_nx, wg_ny, wg_nf, wg_avg;
    double wg_x0, wg_y0, wg_dx, wg_dy, wg_t0, wg_fdt;
    double *wg_ua, *wg_va;            /* (frames, ny, nx) each */
    /* hash */
    pstate *tab_p;
    size_t cap, count;
    /* search scratch */
    cand *cbuf;
    int *cell_start, *cell_items, ncells_alloc;
    /* log */
    int32_t *log_i, *log_j, *log_f;
    int nlog, logcap;
    int stats[4];
};

static void *xcalloc(size_t n, size_t sz) {
    void *p = calloc(n ? n : 1, sz);
    if (!p) abort();
    return p;
}

static uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint32_t or_bond_threshold(double frac) {
    if (frac >= 1.0) return 1u << 24;
    if (frac <= 0.0) return 0u;
    return (uint32_t)floor(frac * 16777216.0);
}

int or_bond_keep(int i, int j, uint64_t seed, uint32_t thr24) {
    uint64_t key = mix64(seed ^ (((uint64_t)(uint32_t)i << 32) | (uint64_t)(uint32_t)j));
    return (uint32_t)(key >> 40) < thr24;
}

static size_t hslot(uint64_t key, size_t cap) { return (size_t)(mix64(key) & (cap - 1)); }

static void hash_grow(or_sim *s) {
    size_t ncap = s->cap ? s->cap * 2 : 1024;
    pstate *nt = xcalloc(ncap, sizeof(pstate));
    for (size_t k = 0; k < s->cap; ++k) {
        if (!s->tab_p[k].used) continue;
        size_t h = hslot(s->tab_p[k].key, ncap);
        while (nt[h].used) h = (h + 1) & (ncap - 1);
        nt[h] = s->tab_p[k];
    }
    free(s->tab_p);
    s->tab_p = nt;
    s->cap = ncap;
}

static pstate *pfind(or_sim *s, int i, int j) {
    if (!s->cap) return NULL;
    uint64_t key = ((uint64_t)i << 32) | (uint64_t)j;
    size_t h = hslot(key, s->cap);
    while (s->tab_p[h].used) {
        if (s->tab_p[h].key == key) return &s->tab_p[h];
        h = (h + 1) & (s->cap - 1);
    }
    return NULL;
}

static pstate *pget(or_sim *s, int i, int j) {
    pstate *e = pfind(s, i, j);
    if (e) return e;
    if ((s->count + 1) * 2 > s->cap) hash_grow(s);
    uint64_t key = ((uint64_t)i << 32) | (uint64_t)j;
    size_t h = hslot(key, s->cap);
    while (s->tab_p[h].used) h = (h + 1) & (s->cap - 1);
    memset(&s->tab_p[h], 0, sizeof(pstate));
    s->tab_p[h].used = 1;
    s->tab_p[h].key = key;
    s->count++;
    return &s->tab_p[h];
}

/* ------------------------------------------------------------------ */
/* physics                                                              */
/* ------------------------------------------------------------------ */

/* floe_properties, physics.f90:179-197 */
static void floe_properties(or_sim *s, int i) {
    const or_params *p = &s->p;
    double r = s->r[i], h = s->h[i];
    s->mass[i] = p->rhoice * s->pi * h * (r * r);
    s->inertia[i] = 0.5 * s->mass[i] * (r * r);
    s->hfa[i] = h * (p->rhowater - p->rhoice) / p->rhowater;
    s->hfw[i] = h * p->rhoice / p->rhowater;
}

/* fit, contact_forces.f90:247-265 */
static double fit(double xi) {
    const double p1 = 9117e-4, p2 = 2722e-4, p3 = 3324e-6, q1 = 1524e-3, q2 = 3159e-5;
    return (p1 * (xi * xi) + p2 * xi + p3) / ((xi * xi) + q1 * xi + q2);
}

/* the fit function alone, for the closed-form tests */
double or_fit(double xi) { return fit(xi); }

/* heaviside, forcing.f90:226-236 */
static double heaviside(double x) { return copysign(0.5, x) + 0.5; }

/* S_shelter, forcing.f90:238-270 (0.18 is a REAL*4 literal) */
static double S_shelter(double hfi, double hfj, double deltan, double cosa, double sina,
                        double uf, double vf) {
    const double c018 = (double)0.18f;
    double wnorm = sqrt(uf * uf + vf * vf) + 1e-20;
    double costheta = (uf * cosa + vf * sina) / wnorm;
    double Hd = heaviside(deltan);
    double D = 1.0 - (1.0 - Hd) * (1.0 - exp(-(c018 * fabs(deltan / costheta)) / hfj))
             - Hd * heaviside(hfi - hfj) * (1.0 - hfj / hfi);
    double G = heaviside(costheta) * costheta;
    return 1.0 - D * G;
}

/* log_profile, forcing.f90:209-224 */
static double log_profile(const or_sim *s, double hf) {
    double t = log(hf / s->p.z0w) / s->log10z0;
    return t * t;
}

typedef struct pairkin {
    double dist, cosa, sina, omegarel, veln, veltb, velt, deltan;
} pairkin;

/* rel_pos_vel, physics.f90:73-114 (alpha is never read and is omitted) */
static void rel_pos_vel(const or_sim *s, int j, int i, pairkin *k) {
    double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j];
    k->dist = sqrt(dx * dx + dy * dy);
    k->cosa = (s->x[j] - s->x[i]) / k->dist;
    k->sina = (s->y[j] - s->y[i]) / k->dist;
    k->omegarel = s->omega[j] - s->omega[i];
    k->veln = (s->u[j] - s->u[i]) * k->cosa + (s->v[j] - s->v[i]) * k->sina;
    k->veltb = -(s->u[j] - s->u[i]) * k->sina + (s->v[j] - s->v[i]) * k->cosa;
    k->velt = k->veltb - (s->omega[i] * s->r[i] + s->omega[j] * s->r[j]);
    k->deltan = s->r[i] + s->r[j] - k->dist;
}

static double delt_ridge_of(const or_sim *s, int j, int i, double dist) {
    double ri = s->r[i], rj = s->r[j];
    double t = (dist * dist - rj * rj + ri * ri) / (2.0 * dist);
    return 2.0 * sqrt(ri * ri - t * t);
}

/* update_shape, ridging.f90:36-99; returns the floe that was reshaped (-1 none) */
static int update_shape(or_sim *s, int j, int i, double dist, double *delt_ridge) {
    double hi = s->h[i], hj = s->h[j];
    double hmin = fmin(hi, hj);
    double ri = s->r[i], rj = s->r[j];
    double dij = (dist * dist - rj * rj + ri * ri) / (2.0 * dist);
    double dji = (dist * dist - ri * ri + rj * rj) / (2.0 * dist);
    double area = ri * ri * acos(dij / ri) - dij * *delt_ridge / 2.0
                + rj * rj * acos(dji / rj) - dji * *delt_ridge / 2.0;
    double vol = fmax(area * hmin / 2.0, 0.0);
    int who = -1;
    if (hmin == hi) who = i;
    else if (hmin == hj) who = j;
    if (who >= 0) {
        double rw = s->r[who], hw = s->h[who];
        double dh = vol / (s->pi * (rw * rw));
        s->r[who] = rw * sqrt(hw / (hw + dh));
        s->h[who] = hw + dh;
        floe_properties(s, who);
    }
    *delt_ridge = delt_ridge_of(s, j, i, dist);
    return who;
}

/* coulomb, physics.f90:1-33 */
static void coulomb(const or_sim *s, double velt, double ktc, double gamt, double *fcn,
                    double *fct, double *fcr, double *deltat) {
    const or_params *p = &s->p;
    if (fabs(*fct) > p->friction_coeff * fabs(*fcn)) {
        *fct = -(p->friction_coeff * fabs(*fcn) * copysign(1.0, velt));
        *fcr = *fct;
        if (p->slipping) *deltat = (*fct + gamt * velt) / ktc;
    } else {
        *fcr = *fct;
    }
}

/* contact_forces, contact_forces.f90:1-87, with plastic_contact
 * (ridging.f90:1-33); returns the reshaped floe or -1. */
static int contact_forces(or_sim *s, int j, int i, const pairkin *k, pstate *e, int bonded,
                          double *fcn, double *fct, double *fcr, double *mcc,
                          double *delt_ridge) {
    const or_params *p = &s->p;
    e->deltat = -k->velt * p->dt + e->deltat;
    *delt_ridge = delt_ridge_of(s, j, i, k->dist);
    e->thetarelc = -k->omegarel * p->dt + e->thetarelc;
    double m_redu = s->mass[i] * s->mass[j] / (s->mass[i] + s->mass[j]);
    double r_redu = s->r[i] * s->r[j] / (s->r[i] + s->r[j]);
    double hmin = fmin(s->h[i], s->h[j]);
    double knc = s->pi * s->ec * hmin * fit(k->deltan * r_redu / (2.0 * (hmin * hmin)));
    double ktc = 6.0 * s->gc / s->ec * knc;
    double dr2 = *delt_ridge * *delt_ridge;
    double krc = knc * dr2 / 12.0;
    double gamn = -(s->beta * sqrt(4.0 * knc * m_redu));
    double gamt = -(2.0 * s->beta * sqrt(2.0 / 3.0 * ktc * m_redu));
    double gamr = gamn * dr2 / 12.0;
    *fcn = fmax(knc * k->deltan - gamn * k->veln, 0.0);
    *fct = ktc * e->deltat - gamt * k->velt;
    int who = -1;
    if (p->ridging) {
        if (p->sigmanc_crit * hmin <= *fcn / *delt_ridge / hmin) {
            /* plastic_contact, ridging.f90:1-33 */
            double h2 = hmin * hmin;
            double kp = p->sigmanc_crit * h2 * *delt_ridge / k->deltan;
            ktc = 6.0 * s->gc / s->ec * kp;
            krc = kp * dr2 / 12.0;
            gamn = -(s->beta * sqrt(4.0 * kp * m_redu));
            gamt = -(2.0 * s->beta * sqrt(2.0 / 3.0 * ktc * m_redu));
            gamr = gamn * dr2 / 12.0;
            *fcn = fmin(*fcn, p->sigmanc_crit * h2 * *delt_ridge);
            *fct = ktc * e->deltat - gamt * k->velt;
            who = update_shape(s, j, i, k->dist, delt_ridge);
        }
    }
    coulomb(s, k->velt, ktc, gamt, fcn, fct, fcr, &e->deltat);
    if (!bonded) {
        double mrolling = krc * e->thetarelc - gamr * k->omegarel;
        if (fabs(e->thetarelc) > 2.0 * fabs(*fcn) / knc / *delt_ridge) mrolling = 0.0;
        *mcc = mrolling;
    }
    return who;
}

/* bond_properties, bonds.f90:237-268 (rb**3d0 folds to ((rb*rb)*rb)) */
static void bond_properties(const or_sim *s, int j, int i, double dist, pstate *e) {
    const or_params *p = &s->p;
    e->rb = p->lambda_rb * fmin(s->r[i], s->r[j]);
    e->hb = (s->h[i] + s->h[j]) / 2.0;
    e->lb = dist;
    e->sb = 2.0 * e->rb * e->hb;
    e->ib = 2.0 / 3.0 * e->hb * (e->rb * e->rb * e->rb);
}

/* bond_forces_euler, bonds.f90:74-161 */
static void bond_forces_euler(const or_sim *s, int j, int i, const pairkin *k, pstate *e) {
    const or_params *p = &s->p;
    e->deltanb = k->veln * p->dt + e->deltanb;
    e->deltatb = k->veltb * p->dt + e->deltatb;
    e->trb_ji = s->omega[i] * p->dt + e->trb_ji;
    e->trb_ij = s->omega[j] * p->dt + e->trb_ij;
    double m_redu = s->mass[i] * s->mass[j] / (s->mass[i] + s->mass[j]);
    double i_redu = s->inertia[i] * s->inertia[j] / (s->inertia[i] + s->inertia[j]);
    double L = e->lb;
    double EA = (1.0 - e->damage) * p->eb * e->sb;
    double EI = (1.0 - e->damage) * p->eb * e->ib;
    double L2 = L * L, L3 = L * L * L;
    double k_axial = EA / L;
    double k_shear1 = 12.0 * EI / L3;
    double k_shear2 = 6.0 * EI / L2;
    double k_rot1 = 4.0 * EI / L;
    double k_rot2 = 2.0 * EI / L;
    double eta_i = 2.0 * sqrt(EI * i_redu / L);
    double eta_a = 2.0 * sqrt(EA * m_redu * L);
    double gam_axial = eta_a / L;
    double gam_shear1 = 12.0 * eta_i / L3;
    double gam_shear2 = 6.0 * eta_i / L2;
    double gam_rot1 = 4.0 * eta_i / L;
    double gam_rot2 = 2.0 * eta_i / L;
    double oi = s->omega[i], oj = s->omega[j];
    e->fbn = k_axial * e->deltanb + gam_axial * k->veln;
    e->fbt = k_shear1 * e->deltatb + k_shear2 * (e->trb_ji + e->trb_ij)
           + gam_shear1 * k->veltb + gam_shear2 * (oi + oj);
    e->mbb_ji = k_rot1 * e->trb_ji + k_rot2 * e->trb_ij - k_shear2 * e->deltatb
              + gam_rot1 * oi + gam_rot2 * oj - gam_shear2 * k->veltb;
    e->mbb_ij = k_rot2 * e->trb_ji + k_rot1 * e->trb_ij - k_shear2 * e->deltatb
              + gam_rot2 * oi + gam_rot1 * oj - gam_shear2 * k->veltb;
}

/* bond_breaking, bonds.f90:164-212 */
static void bond_breaking(const or_sim *s, pstate *e) {
    const or_params *p = &s->p;
    double taub = fabs(e->fbt) / e->sb;
    double sigmatb = -e->fbn / e->sb + fabs(e->mbb_ji) * e->rb / e->ib;
    double sigmacb = e->fbn / e->sb + fabs(e->mbb_ji) * e->rb / e->ib;
    double t = taub / (p->tau_crit * e->hb);
    double phi = t * t;
    if (sigmacb > 0) { double c = sigmacb / (p->sigmacb_crit * e->hb); phi = phi + c * c; }
    if (sigmatb > 0) { double c = sigmatb / (p->sigmatb_crit * e->hb); phi = phi + c * c; }
    double psi = fmin(1.0, 1.0 / sqrt(phi));
    e->damage = e->damage + (1.0 - psi) * (1.0 - e->damage) * p->dt / p->dtd
              - e->damage * p->dt / p->dth;
    if (e->damage >= p->dmax) {
        e->bond = 0;
        e->fbn = 0.0;
        e->fbt = 0.0;
        e->mbb_ji = 0.0;
        e->damage = 1.0;
    }
}

/* coulomb_bc, physics.f90:36-70 */
static void coulomb_bc(or_sim *s, int i, double velt_bc, double ktc, double gamt, double *deltat_bc) {
    const or_params *p = &s->p;
    if (fabs(s->ft_bc[i]) > p->friction_coeff * fabs(s->fn_bc[i])) {
        s->ft_bc[i] = -(p->friction_coeff * fabs(s->fn_bc[i]) * copysign(1.0, velt_bc));
        s->fr_bc[i] = s->ft_bc[i];
        if (p->slipping) *deltat_bc = (s->ft_bc[i] + gamt * velt_bc) / ktc;
    } else {
        s->fr_bc[i] = s->ft_bc[i];
    }
}

/* contact_bc, contact_forces.f90:90-244, with plastic_contact_bc and
 * update_shape_bc (ridging.f90:102-179).  Returns 1 if floe i was reshaped. */
static int contact_bc(or_sim *s, int i, int dir1, int dir2, int bd) {
    const or_params *p = &s->p;
    double c = (double)(dir1 * (dir2 - (1 - dir2)));
    double sn = (double)((1 - dir1) * (dir2 - (1 - dir2)));
    double d1 = (double)dir1, d2 = (double)dir2, e1 = (double)(1 - dir1), e2 = (double)(1 - dir2);
    double ui = s->u[i], vi = s->v[i], oi = s->omega[i];
    double veln_bc = -ui * c - vi * sn;
    double velt_bc = ui * sn - vi * c - oi * s->r[i];
    double ri = s->r[i];
    double deltan_bc = (e2 * (ri - s->x[i]) + d2 * (ri + s->x[i] - p->nx)) * d1
                     + (e2 * (ri - s->y[i]) + d2 * (ri + s->y[i] - p->ny)) * e1;
    double deltat_bc = 0.0;
    double t = ri - deltan_bc;
    s->delt_ridge_bc[i] = 2.0 * sqrt(ri * ri - t * t);
    if (bd == 1) {
        s->deltat_bc1[i] = -velt_bc * p->dt + s->deltat_bc1[i];
        deltat_bc = s->deltat_bc1[i];
    } else if (bd == 2) {
        s->deltat_bc2[i] = -velt_bc * p->dt + s->deltat_bc2[i];
        deltat_bc = s->deltat_bc2[i];
    }
    if (bd == 1) s->theta_bc1[i] = -oi * p->dt + s->theta_bc1[i];
    else if (bd == 2) s->theta_bc2[i] = -oi * p->dt + s->theta_bc2[i];
    double hi = s->h[i];
    double knc = s->pi * s->ec * hi * fit(deltan_bc * ri / (1.0 * 2.0 * (hi * hi)));
    double ktc = 6.0 * s->gc / s->ec * knc;
    double dr2 = s->delt_ridge_bc[i] * s->delt_ridge_bc[i];
    double krc = knc * dr2 / 12.0;
    double gamn = -(s->beta * sqrt(4.0 * knc * s->mass[i] / 1.0));
    double gamt = -(2.0 * s->beta * sqrt(2.0 / 3.0 * ktc * s->mass[i] / 1.0));
    double gamr = gamn * dr2 / 12.0;
    s->fn_bc[i] = fmax(knc * deltan_bc - gamn * veln_bc, 0.0);
    s->ft_bc[i] = ktc * deltat_bc - gamt * velt_bc;
    int reshaped = 0;
    if (p->ridging) {
        if (p->sigmanc_crit * hi <= s->fn_bc[i] / s->delt_ridge_bc[i] / hi) {
            /* plastic_contact_bc, ridging.f90:102-137.  Its dashpots use
             * m(i), the total moment, which reset_forces zeroed for this
             * step (reset.f90:97): they are sqrt(0)-scaled here too. */
            const double m_total_during_loop = 0.0;
            double h2 = hi * hi;
            double kp = p->sigmanc_crit * h2 * s->delt_ridge_bc[i] / deltan_bc;
            ktc = 6.0 * s->gc / s->ec * kp;
            krc = kp * dr2 / 12.0;
            gamn = -(s->beta * sqrt(4.0 * kp * m_total_during_loop / 1.0));
            gamt = -(2.0 * s->beta * sqrt(2.0 / 3.0 * ktc * m_total_during_loop / 1.0));
            gamr = gamn * dr2 / 12.0;
            s->fn_bc[i] = fmin(s->fn_bc[i], p->sigmanc_crit * h2 * s->delt_ridge_bc[i]);
            s->ft_bc[i] = ktc * deltat_bc - gamt * velt_bc;
            /* update_shape_bc, ridging.f90:140-179 */
            double drb = s->delt_ridge_bc[i];
            double area = ri * ri * asin(drb / 2.0 / ri) - drb * (ri - deltan_bc) / 2.0;
            double vol = fmax(area * hi, 0.0);
            double dh = vol / (s->pi * (ri * ri));
            s->r[i] = ri * sqrt(hi / (hi + dh));
            s->h[i] = hi + dh;
            floe_properties(s, i);
            double rn = s->r[i], tn = rn - deltan_bc;
            s->delt_ridge_bc[i] = 2.0 * sqrt(rn * rn - tn * tn);
            reshaped = 1;
        }
    }
    double mrolling_bc;
    if (bd == 1) {
        coulomb_bc(s, i, velt_bc, ktc, gamt, &s->deltat_bc1[i]);
        mrolling_bc = krc * s->theta_bc1[i] - gamr * oi;
        if (fabs(s->theta_bc1[i]) > 2.0 * fabs(s->fn_bc[i]) / knc / s->delt_ridge_bc[i]) mrolling_bc = 0.0;
    } else {
        coulomb_bc(s, i, velt_bc, ktc, gamt, &s->deltat_bc2[i]);
        mrolling_bc = krc * s->theta_bc2[i] - gamr * oi;
        if (fabs(s->theta_bc2[i]) > 2.0 * fabs(s->fn_bc[i]) / knc / s->delt_ridge_bc[i]) mrolling_bc = 0.0;
    }
    s->mc_bc[i] = mrolling_bc;
    return reshaped;
}

/* check_wall, boundaries.f90:100-146 */
static int check_wall(or_sim *s, int i, int dir1, int dir2, int bd) {
    int reshaped = contact_bc(s, i, dir1, dir2, bd);
    double c = (double)(dir1 * (dir2 - (1 - dir2)));
    double sn = (double)((1 - dir1) * (dir2 - (1 - dir2)));
    s->fx_bc[i] = s->fx_bc[i] - s->fn_bc[i] * c + s->fr_bc[i] * sn;
    s->fy_bc[i] = s->fy_bc[i] - s->fn_bc[i] * sn - s->fr_bc[i] * c;
    s->m_bc[i] = s->m_bc[i] - s->mc_bc[i] - s->r[i] * s->ft_bc[i];
    double fnt = sqrt(s->fn_bc[i] * s->fn_bc[i] + s->ft_bc[i] * s->ft_bc[i]);
    s->sigxx_bc[i] = -fnt * s->r[i] * (c * c);
    s->sigyy_bc[i] = -fnt * s->r[i] * (sn * sn);
    s->sigxy_bc[i] = -fnt * s->r[i] * c * sn;
    s->sigyx_bc[i] = -fnt * s->r[i] * sn * c;
    s->ta_bc[i] = s->ta_bc[i] + s->delt_ridge_bc[i] * s->h[i];
    s->p_bc[i] = s->p_bc[i] - s->fn_bc[i] * s->delt_ridge_bc[i] * s->h[i];
    return reshaped;
}

/* reset_boundary, reset.f90:37-69 */
static void reset_boundary(or_sim *s, int i, int bd) {
    if (bd == 1) { s->theta_bc1[i] = 0.0; s->deltat_bc1[i] = 0.0; }
    else if (bd == 2) { s->theta_bc2[i] = 0.0; s->deltat_bc2[i] = 0.0; }
    else { s->theta_bc1[i] = 0.0; s->theta_bc2[i] = 0.0; s->deltat_bc1[i] = 0.0; s->deltat_bc2[i] = 0.0; }
    s->delt_ridge_bc[i] = 0.0;
}

/* verify_bc, boundaries.f90:1-97; returns number of wall reshapes */
static int verify_bc(or_sim *s, int i) {
    const or_params *p = &s->p;
    int t1 = 0, t2 = 0, rs = 0;
    if (s->y[i] - s->r[i] < 0) { t1 = 1; rs += check_wall(s, i, 0, 0, 1); }
    if (s->y[i] + s->r[i] > p->ny) { t1 = 1; rs += check_wall(s, i, 0, 1, 1); }
    if (s->x[i] - s->r[i] < 0) { t2 = 1; rs += check_wall(s, i, 1, 0, 2); }
    if (s->x[i] + s->r[i] > p->nx) { t2 = 1; rs += check_wall(s, i, 1, 1, 2); }
    if (t1 + t2 == 0) reset_boundary(s, i, 0);
    if (t2 == 0) reset_boundary(s, i, 2);
    if (t1 == 0) reset_boundary(s, i, 1);
    return rs;
}

/* forcing, forcing.f90:32-159, and coriolis, forcing.f90:189-206 */
static void forcing(or_sim *s, int i) {
    const or_params *p = &s->p;
    double sca = s->loc_a[i], scw = s->loc_w[i];
    double dua = s->ua_i[i] - s->u[i], dva = s->va_i[i] - s->v[i];
    double duw = s->uw_i[i] - s->u[i], dvw = s->vw_i[i] - s->v[i];
    double avn = sqrt(dua * dua + dva * dva);
    double ovn = sqrt(duw * duw + dvw * dvw);
    double alog = log_profile(s, s->hfa[i]);
    double olog = log_profile(s, s->hfw[i]);
    double r = s->r[i], om = s->omega[i], aom = fabs(om);
    double r2 = r * r;
    double fdax, fday, fsax, fsay, fdwx, fdwy, fswx, fswy;
    if (avn == 0.0) {
        fdax = fday = fsax = fsay = 0.0;
    } else {
        fdax = p->rhoair * p->Cdair * s->hfa[i] * r * dua * avn * alog * sca * s->pi / 2.0;
        fday = p->rhoair * p->Cdair * s->hfa[i] * r * dva * avn * alog * sca * s->pi / 2.0;
        fsax = p->rhoair * s->pi * r2 * p->Csair * dua / avn * (avn * avn + r2 * (aom * aom) / 4.0);
        fsay = p->rhoair * s->pi * r2 * p->Csair * dva / avn * (avn * avn + r2 * (aom * aom) / 4.0);
    }
    if (ovn == 0.0) {
        fdwx = fdwy = fswx = fswy = 0.0;
    } else {
        fdwx = p->rhowater * p->Cdwater * s->hfw[i] * r * duw * ovn * olog * scw * s->pi / 2.0;
        fdwy = p->rhowater * p->Cdwater * s->hfw[i] * r * dvw * ovn * olog * scw * s->pi / 2.0;
        fswx = p->rhowater * s->pi * r2 * p->Cswater * duw / ovn * (ovn * ovn + r2 * (aom * aom) / 4.0);
        fswy = p->rhowater * s->pi * r2 * p->Cswater * dvw / ovn * (ovn * ovn + r2 * (aom * aom) / 4.0);
    }
    s->fax[i] = fdax + fsax;
    s->fay[i] = fday + fsay;
    s->fwx[i] = fdwx + fswx;
    s->fwy[i] = fdwy + fswy;
    double r4 = r * r * r * r, r5 = r * r * r * r * r;
    if (avn < 1e-2) s->ma[i] = -(2.0 * s->pi / 5.0 * r5 * p->rhoair * p->Csair * om * aom);
    else s->ma[i] = -(3.0 / 4.0 * s->pi * p->rhoair * p->Csair * avn * om * r4);
    if (ovn < 1e-3) s->mw[i] = -(2.0 * s->pi / 5.0 * r5 * p->rhowater * p->Cswater * om * aom);
    else s->mw[i] = -(3.0 / 4.0 * s->pi * p->rhowater * p->Cswater * ovn * om * r4);
    s->sigxx_aw[i] = (s->fax[i] + s->fwx[i]) * r;
    s->sigyy_aw[i] = (s->fay[i] + s->fwy[i]) * r;
    s->sigxy_aw[i] = 0.0;
    s->sigyx_aw[i] = 0.0;
    /* coriolis: omega_earth = 7.2921d-5 */
    s->fcorx[i] = s->mass[i] * 2.0 * 7.2921e-5 * s->sinlat * s->v[i];
    s->fcory[i] = -(s->mass[i] * 2.0 * 7.2921e-5 * s->sinlat * s->u[i]);
}

/* ------------------------------------------------------------------ */
/* neighbour search                                                     */
/* ------------------------------------------------------------------ */
static int cand_cmp(const void *a, const void *b) {
    const cand *x = a, *y = b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->j > y->j) - (x->j < y->j);
}

/* grid of cell size >= max query radius, rebuilt each step */
typedef struct grid { double x0, y0, cs; int gx, gy; } grid;

static void build_grid(or_sim *s, grid *g) {
    int n = s->n;
    double xmin = s->x[0], xmax = s->x[0], ymin = s->y[0], ymax = s->y[0], rmax = 0;
    for (int i = 0; i < n; ++i) {
        if (s->x[i] < xmin) xmin = s->x[i];
        if (s->x[i] > xmax) xmax = s->x[i];
        if (s->y[i] < ymin) ymin = s->y[i];
        if (s->y[i] > ymax) ymax = s->y[i];
        if (s->r[i] > rmax) rmax = s->r[i];
    }
    g->cs = (rmax + s->p.rtree) * 1.0000001 + 1e-9;
    g->x0 = xmin;
    g->y0 = ymin;
    g->gx = (int)((xmax - xmin) / g->cs) + 1;
    g->gy = (int)((ymax - ymin) / g->cs) + 1;
    long nc = (long)g->gx * g->gy;
    if (nc > 16L * n + 1024) { g->gx = g->gy = 1; nc = 1; g->cs = INFINITY; }
    if (nc + 1 > s->ncells_alloc) {
        free(s->cell_start);
        s->ncells_alloc = (int)nc + 1;
        s->cell_start = xcalloc(s->ncells_alloc, sizeof(int));
    }
    memset(s->cell_start, 0, sizeof(int) * (nc + 1));
    int *cellof = xcalloc(n, sizeof(int));
    for (int i = 0; i < n; ++i) {
        int cx = g->gx == 1 ? 0 : (int)((s->x[i] - g->x0) / g->cs);
        int cy = g->gy == 1 ? 0 : (int)((s->y[i] - g->y0) / g->cs);
        if (cx >= g->gx) cx = g->gx - 1;
        if (cy >= g->gy) cy = g->gy - 1;
        cellof[i] = cy * g->gx + cx;
        s->cell_start[cellof[i] + 1]++;
    }
    for (long c = 0; c < nc; ++c) s->cell_start[c + 1] += s->cell_start[c];
    int *fill = xcalloc(nc + 1, sizeof(int));
    memcpy(fill, s->cell_start, sizeof(int) * (nc + 1));
    for (int i = 0; i < n; ++i) s->cell_items[fill[cellof[i]]++] = i;
    free(fill);
    free(cellof);
}

/* Candidates of i: {j : dist <= r(i)+rtree} sorted by (dist, j), including i
 * itself; stepper.f90:77-78 (coretran kNearest with a radius). */
static int candidates(or_sim *s, const grid *g, int i) {
    double rad = s->r[i] + s->p.rtree;
    int cx = g->gx == 1 ? 0 : (int)((s->x[i] - g->x0) / g->cs);
    int cy = g->gy == 1 ? 0 : (int)((s->y[i] - g->y0) / g->cs);
    if (cx >= g->gx) cx = g->gx - 1;
    if (cy >= g->gy) cy = g->gy - 1;
    int cnt = 0;
    for (int b = cy - 1; b <= cy + 1; ++b) {
        if (b < 0 || b >= g->gy) continue;
        for (int a = cx - 1; a <= cx + 1; ++a) {
            if (a < 0 || a >= g->gx) continue;
            int c = b * g->gx + a;
            for (int q = s->cell_start[c]; q < s->cell_start[c + 1]; ++q) {
                int j = s->cell_items[q];
                double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j];
                double d = sqrt(dx * dx + dy * dy);
                if (d <= rad) { s->cbuf[cnt].d = d; s->cbuf[cnt].j = j; cnt++; }
            }
        }
    }
    qsort(s->cbuf, cnt, sizeof(cand), cand_cmp);
    return cnt;
}

/* ------------------------------------------------------------------ */
/* public API                                                           */
/* ------------------------------------------------------------------ */
or_sim *or_create(const or_params *p, int n) {
    or_sim *s = xcalloc(1, sizeof(or_sim));
    s->p = *p;
    s->n = n;
    s->pi = 4.0 * atan(1.0);
    s->ec = p->e_modul / (2.0 * (1.0 - p->poiss_ratio * p->poiss_ratio));
    s->gc = p->e_modul / (4.0 * (1.0 - p->poiss_ratio) * (2.0 + p->poiss_ratio));
    double lr = log(p->rest_coeff);
    s->beta = 10.0 * lr / sqrt(lr * lr + s->pi * s->pi);
    s->sinlat = sin(p->lat);
    s->log10z0 = log(10.0 / p->z0w);
    double **arrs[] = {&s->x, &s->y, &s->r, &s->h, &s->hfa, &s->hfw, &s->mass, &s->inertia,
        &s->u, &s->v, &s->theta, &s->omega, &s->fcx, &s->fcy, &s->mc, &s->fbx, &s->fby, &s->mb,
        &s->fax, &s->fay, &s->fwx, &s->fwy, &s->ma, &s->mw, &s->fcorx, &s->fcory,
        &s->fn_bc, &s->ft_bc, &s->fr_bc, &s->fx_bc, &s->fy_bc, &s->mc_bc, &s->m_bc,
        &s->deltat_bc1, &s->deltat_bc2, &s->theta_bc1, &s->theta_bc2, &s->delt_ridge_bc,
        &s->tfx, &s->tfy, &s->m, &s->loc_a, &s->loc_w, &s->thr_a, &s->thr_w,
        &s->ua_i, &s->va_i, &s->uw_i, &s->vw_i, &s->sigxx, &s->sigyy, &s->sigxy, &s->sigyx,
        &s->sigxx_bc, &s->sigyy_bc, &s->sigxy_bc, &s->sigyx_bc,
        &s->sigxx_aw, &s->sigyy_aw, &s->sigxy_aw, &s->sigyx_aw,
        &s->tac, &s->tab, &s->pc, &s->pb, &s->tp, &s->ta_bc, &s->p_bc,
        &s->tsig[0], &s->tsig[1], &s->tsig[2], &s->tsig[3]};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); ++k) *arrs[k] = xcalloc(n, sizeof(double));
    s->cbuf = xcalloc(n + 1, sizeof(cand));
    s->cell_items = xcalloc(n, sizeof(int));
    s->logcap = 0;
    return s;
}

void or_destroy(or_sim *s) {
    if (!s) return;
    double *arrs[] = {s->x, s->y, s->r, s->h, s->hfa, s->hfw, s->mass, s->inertia,
        s->u, s->v, s->theta, s->omega, s->fcx, s->fcy, s->mc, s->fbx, s->fby, s->mb,
        s->fax, s->fay, s->fwx, s->fwy, s->ma, s->mw, s->fcorx, s->fcory,
        s->fn_bc, s->ft_bc, s->fr_bc, s->fx_bc, s->fy_bc, s->mc_bc, s->m_bc,
        s->deltat_bc1, s->deltat_bc2, s->theta_bc1, s->theta_bc2, s->delt_ridge_bc,
        s->tfx, s->tfy, s->m, s->loc_a, s->loc_w, s->thr_a, s->thr_w,
        s->ua_i, s->va_i, s->uw_i, s->vw_i, s->sigxx, s->sigyy, s->sigxy, s->sigyx,
        s->sigxx_bc, s->sigyy_bc, s->sigxy_bc, s->sigyx_bc,
        s->sigxx_aw, s->sigyy_aw, s->sigxy_aw, s->sigyx_aw,
        s->tac, s->tab, s->pc, s->pb, s->tp, s->ta_bc, s->p_bc,
        s->tsig[0], s->tsig[1], s->tsig[2], s->tsig[3]};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); ++k) free(arrs[k]);
    free(s->wg_ua); free(s->wg_va);
    free(s->tab_p); free(s->cbuf); free(s->cell_start); free(s->cell_items);
    free(s->log_i); free(s->log_j); free(s->log_f);
    free(s);
}

/* ini_get.f90:145-247: state read, pair arrays zeroed, floe_properties */
void or_set_state(or_sim *s, const double *x, const double *y, const double *r,
                  const double *h, const double *theta, const double *omega,
                  const double *u, const double *v) {
    int n = s->n;
    memcpy(s->x, x, n * sizeof(double)); memcpy(s->y, y, n * sizeof(double));
    memcpy(s->r, r, n * sizeof(double)); memcpy(s->h, h, n * sizeof(double));
    memcpy(s->theta, theta, n * sizeof(double)); memcpy(s->omega, omega, n * sizeof(double));
    memcpy(s->u, u, n * sizeof(double)); memcpy(s->v, v, n * sizeof(double));
    for (int i = 0; i < n; ++i) {
        s->theta_bc1[i] = s->theta_bc2[i] = s->deltat_bc1[i] = s->deltat_bc2[i] = 0.0;
        s->loc_a[i] = s->loc_w[i] = 0.0;    /* COMMON blocks start zeroed */
        floe_properties(s, i);
    }
    free(s->tab_p);
    s->tab_p = NULL;
    s->cap = s->count = 0;
}

void or_set_wind_frames(or_sim *s, int nxn, int nyn, double x0, double y0, double dx,
                        double dy, int nframes, double t0, double frame_dt, int avg,
                        const double *ua, const double *va) {
    size_t cnt = (size_t)nframes * nxn * nyn;
    free(s->wg_ua); free(s->wg_va);
    s->wg_nx = nxn; s->wg_ny = nyn; s->wg_x0 = x0; s->wg_y0 = y0; s->wg_dx = dx; s->wg_dy = dy;
    s->wg_nf = nframes; s->wg_avg = avg ? 1 : 0;
    s->wg_t0 = t0; s->wg_fdt = nframes > 1 ? frame_dt : 1.0;
    s->wg_ua = xcalloc(cnt, sizeof(double));
    s->wg_va = xcalloc(cnt, sizeof(double));
    memcpy(s->wg_ua, ua, sizeof(double) * cnt);
    memcpy(s->wg_va, va, sizeof(double) * cnt);
}

void or_set_wind_grid(or_sim *s, int nxn, int nyn, double x0, double y0,
                      double dx, double dy, const double *ua, const double *va) {
    or_set_wind_frames(s, nxn, nyn, x0, y0, dx, dy, 1, 0.0, 1.0, 0, ua, va);
}

/* bilinear sample of frame f; identical expression to
 * godar_amd.generators.sample_grid */
static void wind_bilinear(const or_sim *s, int f, double x, double y, double *ua, double *va) {
    double fx = (x - s->wg_x0) / s->wg_dx, fy = (y - s->wg_y0) / s->wg_dy;
    fx = fmin(fmax(fx, 0.0), (double)s->wg_nx - 1.000001);
    fy = fmin(fmax(fy, 0.0), (double)s->wg_ny - 1.000001);
    int a = (int)floor(fx), b = (int)floor(fy);
    double tx = fx - a, ty = fy - b;
    size_t base = (size_t)f * s->wg_nx * s->wg_ny;
    const double *g[2] = {s->wg_ua + base, s->wg_va + base};
    double out[2];
    for (int c = 0; c < 2; ++c) {
        double g00 = g[c][b * s->wg_nx + a], g10 = g[c][b * s->wg_nx + a + 1];
        double g01 = g[c][(b + 1) * s->wg_nx + a], g11 = g[c][(b + 1) * s->wg_nx + a + 1];
        double lo = g00 + (g10 - g00) * tx;
        double hi = g01 + (g11 - g01) * tx;
        out[c] = lo + (hi - lo) * ty;
    }
    *ua = out[0];
    *va = out[1];
}

/* the field at (x, y): frame f0, or linear in time towards frame f0 + 1 */
static void wind_point(const or_sim *s, int f0, double w, double x, double y, double *ua,
                       double *va) {
    wind_bilinear(s, f0, x, y, ua, va);
    if (w > 0.0) {
        double ub, vb;
        wind_bilinear(s, f0 + 1, x, y, &ub, &vb);
        *ua = *ua + (ub - *ua) * w;
        *va = *va + (vb - *va) * w;
    }
}

/* Extension of winds_currents (forcing.f90:1-29), which plans a gridded
 * field averaged over the floe's area: the field at model time t (linear
 * between frames), at the centre or averaged over the disc of radius r by
 * the 7-point rule of degree 5 for the disc (centre weight 1/4, six points
 * at radius r sqrt(2/3), angles k 60 degrees, weight 1/8 each). */
static void sample_wind(const or_sim *s, double x, double y, double r, double t, double *ua,
                        double *va) {
    int f0 = 0;
    double w = 0.0;
    if (s->wg_nf > 1) {
        double sf = (t - s->wg_t0) / s->wg_fdt;
        if (sf >= (double)(s->wg_nf - 1)) {
            f0 = s->wg_nf - 1;
        } else if (sf > 0.0) {
            f0 = (int)floor(sf);
            w = sf - (double)f0;
        }
    }
    wind_point(s, f0, w, x, y, ua, va);
    if (s->wg_avg && r > 0.0) {
        const double s3 = 0x1.bb67ae8584caap-1;      /* sqrt(3)/2 */
        const double rho = r * 0x1.a20bd700c2c3ep-1;  /* r sqrt(2/3) */
        const double cs[6][2] = {{1.0, 0.0}, {0.5, s3}, {-0.5, s3}, {-1.0, 0.0}, {-0.5, -s3}, {0.5, -s3}};
        double su = 0.25 * *ua, sv = 0.25 * *va;
        for (int q = 0; q < 6; ++q) {
            double pu, pv;
            wind_point(s, f0, w, x + rho * cs[q][0], y + rho * cs[q][1], &pu, &pv);
            su = su + 0.125 * pu;
            sv = sv + 0.125 * pv;
        }
        *ua = su;
        *va = sv;
    }
}

void or_get_winds(or_sim *s, double *ua, double *va) {
    memcpy(ua, s->ua_i, sizeof(double) * s->n);
    memcpy(va, s->va_i, sizeof(double) * s->n);
}

static void log_push(or_sim *s, int i, int j, int f) {
    if (s->nlog >= s->logcap) {
        s->logcap = s->logcap ? s->logcap * 2 : 4096;
        s->log_i = realloc(s->log_i, sizeof(int32_t) * s->logcap);
        s->log_j = realloc(s->log_j, sizeof(int32_t) * s->logcap);
        s->log_f = realloc(s->log_f, sizeof(int32_t) * s->logcap);
        if (!s->log_i || !s->log_j || !s->log_f) abort();
    }
    s->log_i[s->nlog] = i; s->log_j[s->nlog] = j; s->log_f[s->nlog] = f;
    s->nlog++;
}

/* One step: stepper.f90:43-343 for one rank and one thread. */
int or_step(or_sim *s, int tstep) {
    const or_params *p = &s->p;
    int n = s->n;
    uint32_t thr24 = or_bond_threshold(p->bond_frac);
    s->nlog = 0;
    memset(s->stats, 0, sizeof(s->stats));
    /* reset_forces, reset.f90:73-178 */
    for (int i = 0; i < n; ++i) {
        s->mc[i] = s->mb[i] = s->fcx[i] = s->fcy[i] = s->fbx[i] = s->fby[i] = 0.0;
        s->m[i] = s->tfx[i] = s->tfy[i] = 0.0;
        s->mc_bc[i] = s->m_bc[i] = s->fn_bc[i] = s->ft_bc[i] = s->fx_bc[i] = s->fy_bc[i] = 0.0;
        s->sigxx[i] = s->sigyy[i] = s->sigxy[i] = s->sigyx[i] = 0.0;
        s->sigxx_bc[i] = s->sigyy_bc[i] = s->sigxy_bc[i] = s->sigyx_bc[i] = 0.0;
        s->tp[i] = s->tac[i] = s->tab[i] = s->pc[i] = s->pb[i] = 0.0;
        s->p_bc[i] = s->ta_bc[i] = 0.0;
        s->thr_a[i] = s->thr_w[i] = 1.0;                 /* stepper.f90:40-41 */
        if (p->shelter) s->loc_a[i] = s->loc_w[i] = 1.0;  /* reset_shelter, reset.f90:181-202 */
    }
    grid g;
    build_grid(s, &g);
    /* winds_currents, forcing.f90:1-29; a gridded wind is taken at the
     * start of the step (model time (tstep - 1) dt), with the radii the
     * floes start the step with */
    for (int i = 0; i < n; ++i) {
        if (s->wg_ua) sample_wind(s, s->x[i], s->y[i], s->r[i], (double)(tstep - 1) * p->dt,
                                  &s->ua_i[i], &s->va_i[i]);
        else { s->ua_i[i] = p->ua; s->va_i[i] = p->va; }
        s->uw_i[i] = p->uw;
        s->vw_i[i] = p->vw;
    }
    for (int i = 0; i < n; ++i) {
        int cnt = candidates(s, &g, i);
        /* stepper.f90:81-87: skip the first entry, then pairs with i >= j */
        for (int k = 1; k < cnt; ++k) {
            int j = s->cbuf[k].j;
            if (i >= j) continue;
            pairkin kin;
            rel_pos_vel(s, j, i, &kin);
            pstate *e = pfind(s, i, j);
            int flags = 0;
            /* step-1 bonds, stepper.f90:93-101 (+ bond mask extension) */
            if (tstep == 1 && p->restart != 1 && p->cohesion) {
                if (kin.deltan >= -p->bond_lim && or_bond_keep(i, j, p->bond_seed, thr24)) {
                    e = pget(s, i, j);
                    e->bond = 1;
                    e->damage = 0.0;
                    bond_properties(s, j, i, kin.dist, e);
                }
            }
            double fcn_r = e ? e->fcn : 0.0, fct_r = e ? e->fct : 0.0;
            double fbn_r = e ? e->fbn : 0.0, fbt_r = e ? e->fbt : 0.0;
            if (kin.deltan > 0) {
                flags |= 1;
                s->stats[3]++;
                if (!e) e = pget(s, i, j);
                double delt_ridge;
                int who = contact_forces(s, j, i, &kin, e, e->bond, &e->fcn, &e->fct, &e->fcr,
                                         &e->mcc, &delt_ridge);
                if (who >= 0) { flags |= 4; s->stats[0]++; }
                fcn_r = e->fcn; fct_r = e->fct;
                /* stepper.f90:113-146 */
                s->fcx[i] = s->fcx[i] - e->fcn * kin.cosa + e->fcr * kin.sina;
                s->fcy[i] = s->fcy[i] - e->fcn * kin.sina - e->fcr * kin.cosa;
                s->mc[i] = s->mc[i] - s->r[i] * e->fct - e->mcc;
                s->fcx[j] = s->fcx[j] + e->fcn * kin.cosa - e->fcr * kin.sina;
                s->fcy[j] = s->fcy[j] + e->fcn * kin.sina + e->fcr * kin.cosa;
                s->mc[j] = s->mc[j] - s->r[j] * e->fct + e->mcc;
                double ac = delt_ridge * fmin(s->h[i], s->h[j]);
                s->tac[i] = s->tac[i] + ac;
                s->pc[i] = s->pc[i] - e->fcn * ac;
                s->tac[j] = s->tac[j] + ac;
                s->pc[j] = s->pc[j] - e->fcn * ac;
            } else if (e) {
                /* reset_contact, reset.f90:1-15 */
                e->thetarelc = 0.0;
                e->deltat = 0.0;
            }
            if (e && e->bond == 1) {
                bond_forces_euler(s, j, i, &kin, e);
                bond_breaking(s, e);
                if (e->bond == 1) {
                    flags |= 2;
                    /* stepper.f90:165-198 */
                    s->fbx[i] = s->fbx[i] + e->fbn * kin.cosa - e->fbt * kin.sina;
                    s->fby[i] = s->fby[i] + e->fbn * kin.sina + e->fbt * kin.cosa;
                    s->mb[i] = s->mb[i] - p->lambda_lb * s->r[i] * e->fbt - e->mbb_ji;
                    s->fbx[j] = s->fbx[j] - e->fbn * kin.cosa + e->fbt * kin.sina;
                    s->fby[j] = s->fby[j] - e->fbn * kin.sina - e->fbt * kin.cosa;
                    s->mb[j] = s->mb[j] - p->lambda_lb * s->r[j] * e->fbt - e->mbb_ij;
                    s->tab[i] = s->tab[i] + e->sb;
                    s->pb[i] = s->pb[i] - e->fbn * e->sb;
                    s->tab[j] = s->tab[j] + e->sb;
                    s->pb[j] = s->pb[j] - e->fbn * e->sb;
                }
                fbn_r = e->fbn; fbt_r = e->fbt;
            } else if (e) {
                /* reset_bond, reset.f90:18-34 */
                e->trb_ji = e->trb_ij = e->deltanb = e->deltatb = 0.0;
            }
            if (p->shelter) {
                /* sheltering, forcing.f90:162-186, and stepper.f90:215-223 */
                double ha_ij = S_shelter(s->hfa[i], s->hfa[j], kin.deltan, kin.cosa, kin.sina, s->ua_i[i], s->va_i[i]);
                double hw_ij = S_shelter(s->hfw[i], s->hfw[j], kin.deltan, kin.cosa, kin.sina, s->uw_i[i], s->vw_i[i]);
                double ha_ji = S_shelter(s->hfa[j], s->hfa[i], kin.deltan, -kin.cosa, -kin.sina, s->ua_i[i], s->va_i[i]);
                double hw_ji = S_shelter(s->hfw[j], s->hfw[i], kin.deltan, -kin.cosa, -kin.sina, s->uw_i[i], s->vw_i[i]);
                s->thr_a[i] = fmin(s->thr_a[i], ha_ji);
                s->thr_a[j] = fmin(s->thr_a[j], ha_ij);
                s->thr_w[i] = fmin(s->thr_w[i], hw_ji);
                s->thr_w[j] = fmin(s->thr_w[j], hw_ij);
            }
            /* stress diagnostics, stepper.f90:238-271 (fcn/fct/fbn/fbt may be stale) */
            {
                double S = sqrt(fcn_r * fcn_r + fct_r * fct_r) + sqrt(fbn_r * fbn_r + fbt_r * fbt_r);
                double c = kin.cosa, sn = kin.sina, ri = s->r[i];
                s->sigxx[i] = s->sigxx[i] - S * c * ri * c;
                s->sigyy[i] = s->sigyy[i] - S * sn * ri * sn;
                s->sigxy[i] = s->sigxy[i] - S * sn * ri * c;
                s->sigyx[i] = s->sigyx[i] - S * c * ri * sn;
                s->sigxx[j] = s->sigxx[j] - S * c * ri * c;
                s->sigyy[j] = s->sigyy[j] - S * sn * ri * sn;
                s->sigxy[j] = s->sigxy[j] - S * sn * ri * c;
                s->sigyx[j] = s->sigyx[j] - S * c * ri * sn;
            }
            log_push(s, i, j, flags);
        }
        s->stats[1] += verify_bc(s, i);
    }
    s->stats[2] = s->nlog;
    /* thread reduction and broadcast_shape (identity at one rank) */
    for (int i = 0; i < n; ++i) {
        s->loc_a[i] = fmin(s->loc_a[i], s->thr_a[i]);
        s->loc_w[i] = fmin(s->loc_w[i], s->thr_w[i]);
    }
    for (int i = 0; i < n; ++i) forcing(s, i);
    /* totals, stepper.f90:312-332 */
    for (int i = 0; i < n; ++i) {
        s->tfx[i] = s->fcx[i] + s->fbx[i] + s->fax[i] + s->fwx[i] + s->fcorx[i] + s->fx_bc[i];
        s->tfy[i] = s->fcy[i] + s->fby[i] + s->fay[i] + s->fwy[i] + s->fcory[i] + s->fy_bc[i];
        s->m[i] = s->mc[i] + s->mb[i] + s->ma[i] + s->mw[i] + s->m_bc[i];
        s->tsig[0][i] = s->sigxx[i] + s->sigxx_bc[i] + s->sigxx_aw[i];
        s->tsig[1][i] = s->sigyy[i] + s->sigyy_bc[i] + s->sigyy_aw[i];
        s->tsig[2][i] = s->sigxy[i] + s->sigxy_bc[i] + s->sigxy_aw[i];
        s->tsig[3][i] = s->sigyx[i] + s->sigyx_bc[i] + s->sigyx_aw[i];
        s->tp[i] = (s->tac[i] != 0.0 ? s->pc[i] / s->tac[i] : 0.0)
                 + (s->tab[i] != 0.0 ? s->pb[i] / s->tab[i] : 0.0)
                 + (s->ta_bc[i] != 0.0 ? s->p_bc[i] / s->ta_bc[i] : 0.0);
    }
    /* velocity and position, integration.f90:1-38 */
    for (int i = 0; i < n; ++i) {
        s->u[i] = s->u[i] + (s->tfx[i] / s->mass[i]) * p->dt;
        s->v[i] = s->v[i] + (s->tfy[i] / s->mass[i]) * p->dt;
        s->omega[i] = s->omega[i] + (s->m[i] / s->inertia[i]) * p->dt;
    }
    for (int i = 0; i < n; ++i) {
        s->x[i] = s->x[i] + s->u[i] * p->dt;
        s->y[i] = s->y[i] + s->v[i] * p->dt;
        s->theta[i] = s->theta[i] + s->omega[i] * p->dt;
    }
    return 0;
}

void or_get_state(or_sim *s, double *x, double *y, double *r, double *h,
                  double *theta, double *omega, double *u, double *v,
                  double *mass, double *inertia, double *hfa, double *hfw) {
    size_t b = sizeof(double) * s->n;
    memcpy(x, s->x, b); memcpy(y, s->y, b); memcpy(r, s->r, b); memcpy(h, s->h, b);
    memcpy(theta, s->theta, b); memcpy(omega, s->omega, b); memcpy(u, s->u, b); memcpy(v, s->v, b);
    memcpy(mass, s->mass, b); memcpy(inertia, s->inertia, b); memcpy(hfa, s->hfa, b); memcpy(hfw, s->hfw, b);
}

void or_get_forces(or_sim *s, double *f) {
    double *cols[33] = {s->tfx, s->tfy, s->m, s->fcx, s->fcy, s->mc, s->fbx, s->fby, s->mb,
        s->fax, s->fay, s->ma, s->fwx, s->fwy, s->mw, s->fcorx, s->fcory, s->fx_bc, s->fy_bc,
        s->m_bc, s->loc_a, s->loc_w, s->tp, s->tsig[0], s->tsig[1], s->tsig[2], s->tsig[3],
        s->tac, s->pc, s->tab, s->pb, s->ta_bc, s->p_bc};
    for (int c = 0; c < 33; ++c) memcpy(f + (size_t)c * s->n, cols[c], sizeof(double) * s->n);
}

void or_get_walls(or_sim *s, double *w) {
    size_t b = sizeof(double) * s->n;
    memcpy(w, s->deltat_bc1, b); memcpy(w + s->n, s->deltat_bc2, b);
    memcpy(w + 2 * s->n, s->theta_bc1, b); memcpy(w + 3 * s->n, s->theta_bc2, b);
}

int or_get_log(or_sim *s, int cap, int32_t *li, int32_t *lj, int32_t *lf) {
    int c = s->nlog < cap ? s->nlog : cap;
    memcpy(li, s->log_i, sizeof(int32_t) * c);
    memcpy(lj, s->log_j, sizeof(int32_t) * c);
    memcpy(lf, s->log_f, sizeof(int32_t) * c);
    return s->nlog;
}

static int pair_cmp(const void *a, const void *b) {
    const int64_t *x = a, *y = b;
    return (*x > *y) - (*x < *y);
}

int or_get_bonds(or_sim *s, int cap, int32_t *bi, int32_t *bj) {
    int64_t *keys = xcalloc(s->count + 1, sizeof(int64_t));
    int c = 0;
    for (size_t k = 0; k < s->cap; ++k)
        if (s->tab_p[k].used && s->tab_p[k].bond == 1) keys[c++] = (int64_t)s->tab_p[k].key;
    qsort(keys, c, sizeof(int64_t), pair_cmp);
    for (int k = 0; k < c && k < cap; ++k) { bi[k] = (int32_t)(keys[k] >> 32); bj[k] = (int32_t)(keys[k] & 0xffffffff); }
    free(keys);
    return c;
}

void or_get_pairs(or_sim *s, int np, const int32_t *pi, const int32_t *pj, double *out) {
    for (int k = 0; k < np; ++k) {
        pstate *e = pfind(s, pi[k], pj[k]);
        pstate z;
        if (!e) { memset(&z, 0, sizeof(z)); e = &z; }
        double v[18] = {e->deltat, e->thetarelc, (double)e->bond, e->damage, e->deltanb, e->deltatb,
                        e->trb_ji, e->trb_ij, e->rb, e->hb, e->lb, e->sb, e->ib,
                        e->fcn, e->fct, e->fbn, e->fbt, e->mcc};
        for (int c = 0; c < 18; ++c) out[(size_t)c * np + k] = v[c];
    }
}

void or_get_stats(or_sim *s, int32_t *st) {
    for (int k = 0; k < 4; ++k) st[k] = s->stats[k];
}
