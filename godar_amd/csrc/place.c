/*
 * Non-overlapping placement of disks in a walled square box, for the
 * workload generators (godar_amd/generators.py).  Host code, built with gcc
 * into godar_amd/libgodar_place.so.
 *
 * 1. Random sequential addition, largest disk first: each disk gets up to
 *    max_tries uniform proposals inside [r+gap, box-r-gap]^2 and takes the
 *    first that clears every placed disk by `gap`.  Overlap checks use one
 *    cell grid per radius octave (cell = 2 * largest radius of the octave),
 *    so a proposal scans a 3x3 block on each level.
 * 2. Disks that found no free spot take their last proposal; a push-apart
 *    relaxation (each overlapping pair moves apart by a quarter of the
 *    violation plus a quarter of the gap, positions clamped to the box) then
 *    removes every overlap.
 * Deterministic for a given seed (splitmix64 stream).
 * godar_relax_disks runs step 2 alone from caller-given centres.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double cs;          /* cell size */
    double rcap;        /* largest radius on this level */
    int g;              /* cells per axis */
    int *head, *next;
} level_t;

static uint64_t sm_state;
static uint64_t sm_next(void) {
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double sm_unif(void) { return (double)(sm_next() >> 11) * (1.0 / 9007199254740992.0); }

static const double *g_r;
static int cmp_desc(const void *a, const void *b) {
    double ra = g_r[*(const int *)a], rb = g_r[*(const int *)b];
    if (ra > rb) return -1;
    if (ra < rb) return 1;
    return (*(const int *)a > *(const int *)b) - (*(const int *)a < *(const int *)b);
}

static int cell_of(const level_t *L, double p) {
    int c = (int)floor(p / L->cs);
    if (c < 0) c = 0;
    if (c >= L->g) c = L->g - 1;
    return c;
}

static int level_index(double r, double rmax, int nlev) {
    int l = 0;
    double cap = rmax;
    while (l < nlev - 1 && r <= cap * 0.5) { cap *= 0.5; ++l; }
    return l;
}

static void level_insert(level_t *L, int i, double x, double y) {
    int c = cell_of(L, y) * L->g + cell_of(L, x);
    L->next[i] = L->head[c];
    L->head[c] = i;
}

/* 1 if disk (x, y, r) clears every placed disk by gap (skipping `self`) */
static int clear_of(level_t *lev, int nlev, const double *px, const double *py, const double *pr,
                    double x, double y, double r, double gap, int self) {
    for (int l = 0; l < nlev; ++l) {
        level_t *L = &lev[l];
        double reach = r + L->rcap + gap;
        int span = (int)ceil(reach / L->cs);
        int cx = cell_of(L, x), cy = cell_of(L, y);
        for (int b = cy - span; b <= cy + span; ++b) {
            if (b < 0 || b >= L->g) continue;
            for (int a = cx - span; a <= cx + span; ++a) {
                if (a < 0 || a >= L->g) continue;
                for (int j = L->head[b * L->g + a]; j >= 0; j = L->next[j]) {
                    if (j == self) continue;
                    double dx = x - px[j], dy = y - py[j], need = r + pr[j] + gap;
                    if (dx * dx + dy * dy < need * need) return 0;
                }
            }
        }
    }
    return 1;
}

static int relax(int n, const double *r, double box, double gap, int max_relax, level_t *lev,
                 int nlev, const int *lvl, double *x, double *y);

int godar_place_disks(int n, const double *r, double box, double gap, uint64_t seed,
                      int max_tries, int max_relax, double *x, double *y, int *stats) {
    if (n <= 0 || box <= 0) return 1;
    double rmax = 0, rmin = INFINITY;
    for (int i = 0; i < n; ++i) {
        if (!(r[i] > 0)) return 1;
        if (r[i] > rmax) rmax = r[i];
        if (r[i] < rmin) rmin = r[i];
        if (2 * (r[i] + gap) >= box) return 2;
    }
    int nlev = 1;
    while (nlev < 12 && rmax / (1 << nlev) >= rmin) ++nlev;
    level_t lev[12];
    for (int l = 0; l < nlev; ++l) {
        lev[l].rcap = rmax / (double)(1 << l);
        lev[l].cs = 2.0 * lev[l].rcap + gap;
        long g = (long)ceil(box / lev[l].cs);
        if (g < 1) g = 1;
        while (g * g > 8L * n + 64 && g > 1) { g = (g + 1) / 2; lev[l].cs *= 2; }
        lev[l].g = (int)g;
        lev[l].head = malloc(sizeof(int) * g * g);
        lev[l].next = malloc(sizeof(int) * n);
        for (long c = 0; c < g * g; ++c) lev[l].head[c] = -1;
    }
    int *order = malloc(sizeof(int) * n);
    int *lvl = malloc(sizeof(int) * n);
    double *pr = malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) { order[i] = i; pr[i] = r[i]; lvl[i] = level_index(r[i], rmax, nlev); }
    g_r = r;
    qsort(order, n, sizeof(int), cmp_desc);
    sm_state = seed ^ 0x5DEECE66Dull;
    int failed = 0;
    for (int k = 0; k < n; ++k) {
        int i = order[k];
        double lo = r[i] + gap, hi = box - r[i] - gap;
        int ok = 0;
        double px = 0, py = 0;
        for (int t = 0; t < max_tries; ++t) {
            px = lo + (hi - lo) * sm_unif();
            py = lo + (hi - lo) * sm_unif();
            if (clear_of(lev, nlev, x, y, pr, px, py, r[i], gap, -1)) { ok = 1; break; }
        }
        if (!ok) ++failed;
        x[i] = px;
        y[i] = py;
        level_insert(&lev[lvl[i]], i, px, py);
    }
    int it = 0;
    if (failed) it = relax(n, r, box, gap, max_relax, lev, nlev, lvl, x, y);
    if (stats) { stats[0] = failed; stats[1] = it; }
    for (int l = 0; l < nlev; ++l) { free(lev[l].head); free(lev[l].next); }
    free(order); free(lvl); free(pr);
    return (failed && it >= max_relax) ? 3 : 0;
}

/* push-apart relaxation over every overlapping pair; returns the sweeps used
 * (max_relax if overlaps remain) */
static int relax(int n, const double *r, double box, double gap, int max_relax, level_t *lev,
                 int nlev, const int *lvl, double *x, double *y) {
    int it = 0;
    {
        double *dx = calloc(n, sizeof(double)), *dy = calloc(n, sizeof(double));
        for (it = 0; it < max_relax; ++it) {
            for (int l = 0; l < nlev; ++l) {
                long g = lev[l].g;
                for (long c = 0; c < g * g; ++c) lev[l].head[c] = -1;
            }
            for (int i = 0; i < n; ++i) level_insert(&lev[lvl[i]], i, x[i], y[i]);
            memset(dx, 0, sizeof(double) * n);
            memset(dy, 0, sizeof(double) * n);
            int bad = 0;
            for (int i = 0; i < n; ++i) {
                for (int l = 0; l < nlev; ++l) {
                    level_t *L = &lev[l];
                    double reach = r[i] + L->rcap + gap;
                    int span = (int)ceil(reach / L->cs);
                    int cx = cell_of(L, x[i]), cy = cell_of(L, y[i]);
                    for (int b = cy - span; b <= cy + span; ++b) {
                        if (b < 0 || b >= L->g) continue;
                        for (int a = cx - span; a <= cx + span; ++a) {
                            if (a < 0 || a >= L->g) continue;
                            for (int j = L->head[b * L->g + a]; j >= 0; j = L->next[j]) {
                                if (j <= i) continue;
                                double ex = x[j] - x[i], ey = y[j] - y[i];
                                double d = sqrt(ex * ex + ey * ey);
                                double viol = r[i] + r[j] + gap - d;
                                if (viol <= 0) continue;
                                ++bad;
                                if (d <= 0) { ex = 1.0; ey = 0.0; d = 1.0; }
                                double s = (0.25 * viol + 0.25 * gap) / d;
                                dx[j] += s * ex; dy[j] += s * ey;
                                dx[i] -= s * ex; dy[i] -= s * ey;
                            }
                        }
                    }
                }
            }
            if (!bad) break;
            for (int i = 0; i < n; ++i) {
                double lo = r[i] + gap, hi = box - r[i] - gap;
                x[i] = fmin(fmax(x[i] + dx[i], lo), hi);
                y[i] = fmin(fmax(y[i] + dy[i], lo), hi);
            }
        }
        free(dx);
        free(dy);
    }
    return it;
}

static int build_levels(int n, const double *r, double box, double gap, level_t *lev, int **lvl_out) {
    double rmax = 0, rmin = INFINITY;
    for (int i = 0; i < n; ++i) {
        if (r[i] > rmax) rmax = r[i];
        if (r[i] < rmin) rmin = r[i];
    }
    int nlev = 1;
    while (nlev < 12 && rmax / (1 << nlev) >= rmin) ++nlev;
    for (int l = 0; l < nlev; ++l) {
        lev[l].rcap = rmax / (double)(1 << l);
        lev[l].cs = 2.0 * lev[l].rcap + gap;
        long g = (long)ceil(box / lev[l].cs);
        if (g < 1) g = 1;
        while (g * g > 8L * n + 64 && g > 1) { g = (g + 1) / 2; lev[l].cs *= 2; }
        lev[l].g = (int)g;
        lev[l].head = malloc(sizeof(int) * g * g);
        lev[l].next = malloc(sizeof(int) * n);
        for (long c = 0; c < g * g; ++c) lev[l].head[c] = -1;
    }
    int *lvl = malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) lvl[i] = level_index(r[i], rmax, nlev);
    *lvl_out = lvl;
    return nlev;
}

/* Overlap removal only, from the caller's initial centres (e.g. a jittered
 * grid); centres are clamped into [r+gap, box-r-gap] first. */
int godar_relax_disks(int n, const double *r, double box, double gap, int max_relax, double *x,
                      double *y, int *sweeps) {
    if (n <= 0 || box <= 0) return 1;
    for (int i = 0; i < n; ++i) {
        if (!(r[i] > 0)) return 1;
        if (2 * (r[i] + gap) >= box) return 2;
        double lo = r[i] + gap, hi = box - r[i] - gap;
        x[i] = fmin(fmax(x[i], lo), hi);
        y[i] = fmin(fmax(y[i], lo), hi);
    }
    level_t lev[12];
    int *lvl;
    int nlev = build_levels(n, r, box, gap, lev, &lvl);
    int it = relax(n, r, box, gap, max_relax, lev, nlev, lvl, x, y);
    if (sweeps) *sweeps = it;
    for (int l = 0; l < nlev; ++l) { free(lev[l].head); free(lev[l].next); }
    free(lvl);
    return it >= max_relax ? 3 : 0;
}
