// Shared host/device types of the GODAR MI355X step library.
#pragma once
#include <cstdint>

namespace godar {

// Per-member constants derived on the host.  Products that the reference
// forms first in a left-to-right chain (pi*ec, 6*gc/ec, ...) are folded here
// with the same operand order, so the device reproduces the reference's
// rounding.
struct DevParams {
    double dt, rtree, nx, ny;
    double rhoair, rhoice, rhowater;
    double Cdair, Cdwater, Csair, Cswater, z0w;
    double pi, ec, gc, beta;
    double pi_ec;          // pi*ec            (contact_forces.f90:35, :163)
    double k6;             // 6d0*gc/ec        (contact_forces.f90:39)
    double two_beta;       // 2d0*beta         (contact_forces.f90:46)
    double sinlat;         // sin(lat)         (forcing.f90:202)
    double log10z0;        // log(10/z0w)      (forcing.f90:222)
    double friction_coeff, sigmanc_crit;
    double eb, lambda_rb, lambda_lb, sigmatb_crit, sigmacb_crit, tau_crit;
    double bond_lim, dmax, dtd, dth;
    double ua, va, uw, vw;
    // cell grid of this member
    double cell, inv_cell_unused;
    int32_t gx, gy;
    int32_t cell_base;     // first global cell id of this member
    int32_t slipping, cohesion, ridging, shelter, restart;
    uint32_t bond_thr24;
    uint64_t bond_seed;
    int32_t wind_grid;     // -1: uniform ua/va, else index into the wind-grid table
    int32_t dense;         // 1: many floes per cell; one wavefront searches each floe
    int32_t span;          // cells searched on each side (cell >= (max r + rtree + skin) / span)
    int32_t pad2_;
    double skin;           // candidate supersets reach this far beyond the query radius
};

// A gridded wind: nframes frames of nodes (a, b) at (x0 + a dx, y0 + b dy),
// frame f at model time t0 + f fdt (linear in time between frames, the
// first and last frame held outside), each frame ua then va, row-major
// (nyn, nxn), frame f at offset + 2 f nxn nyn in the pool.  avg = 1: the
// floe takes the field averaged over its disc (a 7-point rule), else its
// value at the centre.
struct WindGridDesc {
    int32_t nxn, nyn;
    double x0, y0, dx, dy;
    int64_t offset;
    int32_t nframes, avg;
    double t0, fdt;
};

// Member descriptor (device copy).
struct Member {
    int32_t first;         // first global floe index
    int32_t n;             // floe count
    int32_t tstep;         // reference step number of the next step (1-based)
    int32_t active;
    int32_t tstop;         // the member stops before step tstop (0: no limit)
};

// A member takes part in a step while active and short of its stop step.
__host__ __device__ inline bool member_live(const Member& m) {
    return m.active && (m.tstop <= 0 || m.tstep < m.tstop);
}

// Error/status bits (sticky, device side).
enum : uint32_t {
    ERR_OUTER_CAP  = 1u << 0,
    ERR_INNER_CAP  = 1u << 1,
    ERR_HIST_CAP   = 1u << 2,
    ERR_BOND_CAP   = 1u << 3,
    ERR_REPAIR_CAP = 1u << 4,
    ERR_SEED_CAP   = 1u << 5,
    ERR_NONFINITE  = 1u << 6,
    ERR_SUP_CAP    = 1u << 7,
    // a step launched without the build kernels found a build or refresh
    // due (not a failure: the host runs it again with them)
    ERR_NEED_BUILD = 1u << 8,
};

// Build flag bits (DS::rebuild[0]): a floe has moved more than 0.45 skin
// since the last build, so the next step must build the supersets; a floe
// has moved more than 0.3 skin, so a step that has the build kernels builds
// them now (steps without them carry on).
constexpr int32_t BUILD_DUE = 1, BUILD_SOON = 2;

// Bond record (owner = lower floe index), bonds.f90 state of pair (j,i).
struct BondRec {
    int32_t j;             // partner (global floe index), -1 = empty
    int32_t bond;          // 1 bonded, 0 broken
    double damage, deltanb, deltatb, trb_ji, trb_ij, rb, hb, lb, sb, ib;
};

// Bond terms of one bonded event, written by the owner (lower index) for
// both floes' gathers.
struct BondTerms {
    double fbn_c, fbt_s, fbn_s, fbt_c, lri_fbt, lrj_fbt, mbb_ji, mbb_ij, sb, fbn_sb;
};

// Contact history record (owner = lower floe index): deltat, thetarelc,
// and the pair's last contact forces fcn, fct, which the reference keeps in
// its (n,n) arrays after the floes separate and which the stress
// diagnostics read (stepper.f90:238-271).  A record lives while any of the
// four is non-zero.
struct HistRec {
    int32_t j;             // partner, -1 = empty
    int32_t pad_;
    double deltat, thetarelc, fcn, fct;
};

// Per-event result the gather needs, kept in the repair table for events
// the ordered repair recomputed.  Field names follow the stepper.f90
// accumulation lines (113-146 contact, 165-198 bond, 215-223 shelter).
struct EventOut {
    uint32_t flags;        // EV_* bits
    int32_t ridged;        // 0: floe i reshaped, 1: floe j, -1 none
    // contact terms (valid if EV_CONTACT)
    double fcn_c, fcr_s, fcn_s, fcr_c, ri_fct, rj_fct, mcc, ac, fcn_ac;
    // bond terms (valid if EV_BONDED)
    double fbn_c, fbt_s, fbn_s, fbt_c, lri_fbt, lrj_fbt, mbb_ji, mbb_ij, sb, fbn_sb;
    // shelter coefficients: hsf*(j,i) go to floe i, hsf*(i,j) to floe j
    double hsfa_ji, hsfw_ji, hsfa_ij, hsfw_ij;
    // new props of the reshaped floe
    double nr, nh, nmass, ninertia, nhfa, nhfw;
    // stress diagnostics (stepper.f90:238-271): |(fcn, fct)| (current or
    // last contact) and |(fbn, fbt)| (current bond) of the pair, and r(i)
    // after the event
    double sc, sbd, sri;
};

enum : uint32_t {
    EV_CONTACT = 1u << 0,
    EV_BONDED  = 1u << 1,   // bonded after the event (forces accumulated)
    EV_RIDGED  = 1u << 2,
    EV_REMOVED = 1u << 3,   // no longer a candidate (query radius shrank)
    EV_SHELTER = 1u << 4,
    EV_TABLE   = 1u << 5,   // (evf only) the repair table holds the event: ov is its entry
};

// Wall result of one floe (verify_bc, boundaries.f90:1-97).
struct WallOut {
    uint32_t flags;        // bit0 touched a wall, bit2 reshaped
    int32_t pad_;
    double fx_bc, fy_bc, m_bc, ta_bc, p_bc;
    double deltat_bc1, deltat_bc2, theta_bc1, theta_bc2;
    double nr, nh, nmass, ninertia, nhfa, nhfw;
    double sxx_bc, syy_bc, sxy_bc, syx_bc;   // wall stress of the last wall touched
};

}  // namespace godar
