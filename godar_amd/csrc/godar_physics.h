// Device-side GODAR physics: one pair event, one floe's walls, one floe's
// forcing.  Every expression keeps the reference's operand order (Fortran
// evaluates left to right; integer powers expand left to right as amdflang
// emits them) so that the FP64 instantiation rounds like the reference.
// Build with -ffp-contract=off.
//
// MIXED=true evaluates the force laws (contact, bond, shelter, drag) in FP32
// while geometry, history/strain accumulators, ridging volume bookkeeping
// and integration stay FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "godar_types.h"
#include "cr_math.h"

namespace godar {

template <bool MIXED> struct Prec { using T = double; };
template <> struct Prec<true> { using T = float; };

__device__ __forceinline__ double dmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double dmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float dmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float dmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double gsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float gsqrt(float a) { return sqrtf(a); }
// FP64: correctly rounded (cr_math.h), so that they agree with the host libm
// the reference calls wherever that libm rounds correctly
__device__ __forceinline__ double gexp(double a) { return crm::cr_exp(a); }
__device__ __forceinline__ float gexp(float a) { return expf(a); }
__device__ __forceinline__ double glog(double a) { return crm::cr_log(a); }
__device__ __forceinline__ float glog(float a) { return logf(a); }
// acos/asin of the ridging lens area (FP64 in both precisions): correctly
// rounded in FP64 runs; in mixed precision the device library's (within an
// ulp or two), which the FP32 force laws swamp (r^2 h stays conserved
// whatever the area: r' = r sqrt(h / (h + dh)), h' = h + dh)
template <bool MIXED>
__device__ __forceinline__ double ridge_acos(double a) {
    if constexpr (MIXED) return acos(a);
    else return crm::cr_acos(a);
}
template <bool MIXED>
__device__ __forceinline__ double ridge_asin(double a) {
    if constexpr (MIXED) return asin(a);
    else return crm::cr_asin(a);
}
__device__ __forceinline__ double gabs(double a) { return fabs(a); }
__device__ __forceinline__ float gabs(float a) { return fabsf(a); }
__device__ __forceinline__ double gcopysign(double a, double b) { return copysign(a, b); }
__device__ __forceinline__ float gcopysign(float a, float b) { return copysignf(a, b); }

// splitmix64 finaliser; bond mask of the workload generator (extension)
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline bool bond_keep(int i, int j, uint64_t seed, uint32_t thr24) {
    uint64_t key = mix64(seed ^ (((uint64_t)(uint32_t)i << 32) | (uint64_t)(uint32_t)j));
    return (uint32_t)(key >> 40) < thr24;
}

// Floe properties that ridging can change, plus kinematics.
struct Floe {
    double x, y, u, v, omega;
    double r, h, mass, inertia, hfa, hfw;
};

// floe_properties, physics.f90:179-197
__device__ __forceinline__ void floe_properties(const DevParams& P, Floe& f) {
    f.mass = P.rhoice * P.pi * f.h * (f.r * f.r);
    f.inertia = 0.5 * f.mass * (f.r * f.r);
    f.hfa = f.h * (P.rhowater - P.rhoice) / P.rhowater;
    f.hfw = f.h * P.rhoice / P.rhowater;
}

// fit, contact_forces.f90:247-265
template <typename T>
__device__ __forceinline__ T fit(T xi) {
    const T p1 = T(9117e-4), p2 = T(2722e-4), p3 = T(3324e-6), q1 = T(1524e-3), q2 = T(3159e-5);
    return (p1 * (xi * xi) + p2 * xi + p3) / ((xi * xi) + q1 * xi + q2);
}

// heaviside, forcing.f90:226-236
template <typename T>
__device__ __forceinline__ T heaviside(T x) { return gcopysign(T(0.5), x) + T(0.5); }

// S_shelter, forcing.f90:238-270; 0.18 is a REAL*4 literal in the reference
template <typename T>
__device__ __forceinline__ T S_shelter(T hfi, T hfj, T deltan, T cosa, T sina, T uf, T vf) {
    const T c018 = T((double)0.18f);
    T wnorm = gsqrt(uf * uf + vf * vf) + T(1e-20);
    T costheta = (uf * cosa + vf * sina) / wnorm;
    T Hd = heaviside(deltan);
    T D = T(1) - (T(1) - Hd) * (T(1) - gexp(-(c018 * gabs(deltan / costheta)) / hfj))
        - Hd * heaviside(hfi - hfj) * (T(1) - hfj / hfi);
    T G = heaviside(costheta) * costheta;
    return T(1) - D * G;
}

// min(acc, S_shelter(...)) for the per-floe sheltering minima.  An open-water
// pair (deltan < 0) has Hd = 0 and D = 1 - (1 - e) with e = exp(-0.18 |deltan
// / costheta| / hfj) and |costheta| <= 1; once 0.18 |deltan| / hfj >= 20
// (FP32) or 40 (FP64), e is below half an ulp of 1, 1 - e rounds to 1, D = 0
// and the term is exactly 1.  Every term is at most 1, so such a term leaves
// the minimum unchanged and is not evaluated.
template <typename T>
__device__ __forceinline__ T shelter_min(T acc, T hfi, T hfj, T deltan, T cosa, T sina, T uf, T vf) {
    const T lim = sizeof(T) == 4 ? T(20) : T(40);
    if (deltan < T(0) && T((double)0.18f) * -deltan >= lim * hfj) return acc;
    return dmin(acc, S_shelter(hfi, hfj, deltan, cosa, sina, uf, vf));
}

// Old pair state of one event (contact history + bond record).
struct PairState {
    double deltat, thetarelc;
    double fcn, fct;        // last contact forces (kept after separation)
    int bond;               // 0 none/broken, 1 bonded
    bool has_bond_rec;
    double damage, deltanb, deltatb, trb_ji, trb_ij, rb, hb, lb, sb, ib;
};

__device__ __forceinline__ double delt_ridge_of(double ri, double rj, double dist) {
    double t = (dist * dist - rj * rj + ri * ri) / (2.0 * dist);
    return 2.0 * sqrt(ri * ri - t * t);
}

// Pair kinematics, rel_pos_vel (physics.f90:73-114); alpha is never read.
struct Kin {
    double dist, cosa, sina, omegarel, veln, veltb, velt, deltan;
};

__device__ __forceinline__ void kinematics(const Floe& fi, const Floe& fj, Kin& k) {
    double dx = fi.x - fj.x, dy = fi.y - fj.y;
    k.dist = sqrt(dx * dx + dy * dy);
    k.cosa = (fj.x - fi.x) / k.dist;
    k.sina = (fj.y - fi.y) / k.dist;
    k.omegarel = fj.omega - fi.omega;
    k.veln = (fj.u - fi.u) * k.cosa + (fj.v - fi.v) * k.sina;
    k.veltb = -(fj.u - fi.u) * k.sina + (fj.v - fi.v) * k.cosa;
    k.velt = k.veltb - (fi.omega * fi.r + fj.omega * fj.r);
    k.deltan = fi.r + fj.r - k.dist;
}

// Step-1 bonds, stepper.f90:93-101 (+ the bond mask of the workloads) and
// bond_properties, bonds.f90:237-268.
__device__ __forceinline__ void bond_create(const DevParams& P, int tstep, int gi, int gj,
                                            const Floe& fi, const Floe& fj, const Kin& k,
                                            PairState& ps) {
    if (tstep == 1 && P.restart != 1 && P.cohesion) {
        if (k.deltan >= -P.bond_lim && bond_keep(gi, gj, P.bond_seed, P.bond_thr24)) {
            ps.bond = 1;
            ps.has_bond_rec = true;
            ps.damage = 0.0;
            ps.rb = P.lambda_rb * fmin(fi.r, fj.r);
            ps.hb = (fi.h + fj.h) / 2.0;
            ps.lb = k.dist;
            ps.sb = 2.0 * ps.rb * ps.hb;
            ps.ib = 2.0 / 3.0 * ps.hb * (ps.rb * ps.rb * ps.rb);
        }
    }
}

// Contact of a pair with deltan > 0: contact_forces (contact_forces.f90:1-87)
// with plastic_contact/update_shape (ridging.f90:1-99) and coulomb
// (physics.f90:1-33).  fi/fj are reshaped in place when ridging occurs.
template <bool MIXED>
__device__ __forceinline__ void contact_part(const DevParams& P, Floe& fi, Floe& fj, const Kin& k,
                                             PairState& ps, EventOut& o) {
    using T = typename Prec<MIXED>::T;
    // the switches and limits are requested here, at once, and the plastic
    // and Coulomb tests are evaluated outside their branches, so no
    // parameter load waits behind a test
    const int ridging = P.ridging, slipping = P.slipping;
    const T snc = T(P.sigmanc_crit), fric = T(P.friction_coeff);
    o.flags |= EV_CONTACT;
    const double dist = k.dist;
    ps.deltat = -k.velt * P.dt + ps.deltat;
    double delt_ridge = delt_ridge_of(fi.r, fj.r, dist);
    ps.thetarelc = -k.omegarel * P.dt + ps.thetarelc;
    T m_redu = T(fi.mass) * T(fj.mass) / (T(fi.mass) + T(fj.mass));
    T r_redu = T(fi.r) * T(fj.r) / (T(fi.r) + T(fj.r));
    T hmin = dmin(T(fi.h), T(fj.h));
    T knc = T(P.pi_ec) * hmin * fit(T(k.deltan) * r_redu / (T(2) * (hmin * hmin)));
    T ktc = T(P.k6) * knc;
    T dr = T(delt_ridge);
    T dr2 = dr * dr;
    T krc = knc * dr2 / T(12);
    T gamn = -(T(P.beta) * gsqrt(T(4) * knc * m_redu));
    T gamt = -(T(P.two_beta) * gsqrt(T(2.0 / 3.0) * ktc * m_redu));
    T gamr = gamn * dr2 / T(12);
    T fcn = dmax(knc * T(k.deltan) - gamn * T(k.veln), T(0));
    T fct = ktc * T(ps.deltat) - gamt * T(k.velt);
    const bool plastic = (ridging != 0) & (snc * hmin <= fcn / dr / hmin);
    {
        if (plastic) {
            // plastic_contact, ridging.f90:1-33
            T h2 = hmin * hmin;
            T kp = snc * h2 * dr / T(k.deltan);
            ktc = T(P.k6) * kp;
            krc = kp * dr2 / T(12);
            gamn = -(T(P.beta) * gsqrt(T(4) * kp * m_redu));
            gamt = -(T(P.two_beta) * gsqrt(T(2.0 / 3.0) * ktc * m_redu));
            gamr = gamn * dr2 / T(12);
            fcn = dmin(fcn, snc * h2 * dr);
            fct = ktc * T(ps.deltat) - gamt * T(k.velt);
            // update_shape, ridging.f90:36-99 (FP64 volume bookkeeping)
            double hmin64 = fmin(fi.h, fj.h);
            double ri = fi.r, rj = fj.r;
            double dij = (dist * dist - rj * rj + ri * ri) / (2.0 * dist);
            double dji = (dist * dist - ri * ri + rj * rj) / (2.0 * dist);
            double area = ri * ri * ridge_acos<MIXED>(dij / ri) - dij * delt_ridge / 2.0
                        + rj * rj * ridge_acos<MIXED>(dji / rj) - dji * delt_ridge / 2.0;
            double vol = fmax(area * hmin64 / 2.0, 0.0);
            // the thinner floe is reshaped; its six properties are selected
            // field by field (a pointer to either floe, or a select of whole
            // records, would force both floes out of registers)
            int who = -1;
            if (hmin64 == fi.h) who = 0;
            else if (hmin64 == fj.h) who = 1;
            if (who >= 0) {
                Floe w;
                const double rw = who == 0 ? fi.r : fj.r;
                const double hw = who == 0 ? fi.h : fj.h;
                double dh = vol / (P.pi * (rw * rw));
                w.r = rw * sqrt(hw / (hw + dh));
                w.h = hw + dh;
                floe_properties(P, w);
                const bool wi = who == 0;
                fi.r = wi ? w.r : fi.r;             fj.r = wi ? fj.r : w.r;
                fi.h = wi ? w.h : fi.h;             fj.h = wi ? fj.h : w.h;
                fi.mass = wi ? w.mass : fi.mass;    fj.mass = wi ? fj.mass : w.mass;
                fi.inertia = wi ? w.inertia : fi.inertia;
                fj.inertia = wi ? fj.inertia : w.inertia;
                fi.hfa = wi ? w.hfa : fi.hfa;       fj.hfa = wi ? fj.hfa : w.hfa;
                fi.hfw = wi ? w.hfw : fi.hfw;       fj.hfw = wi ? fj.hfw : w.hfw;
                o.flags |= EV_RIDGED;
                o.ridged = who;
                o.nr = w.r; o.nh = w.h; o.nmass = w.mass;
                o.ninertia = w.inertia; o.nhfa = w.hfa; o.nhfw = w.hfw;
            }
            delt_ridge = delt_ridge_of(fi.r, fj.r, dist);
            dr = T(delt_ridge);
        }
    }
    // coulomb, physics.f90:1-33
    const bool coul = gabs(fct) > fric * gabs(fcn);
    fct = coul ? -(fric * gabs(fcn) * gcopysign(T(1), T(k.velt))) : fct;
    T fcr = fct;
    const double slip_dt = double((fct + gamt * T(k.velt)) / ktc);
    ps.deltat = (coul & (slipping != 0)) ? slip_dt : ps.deltat;
    // rolling moment: bonds only form in step 1, before any contact of the
    // pair, so a bonded pair's mcc is still its initial zero
    T mcc = T(0);
    if (ps.bond == 0) {
        T mrolling = krc * T(ps.thetarelc) - gamr * T(k.omegarel);
        if (gabs(T(ps.thetarelc)) > T(2) * gabs(fcn) / knc / dr) mrolling = T(0);
        mcc = mrolling;
    }
    ps.fcn = double(fcn);
    ps.fct = double(fct);
    o.sc = double(gsqrt(fcn * fcn + fct * fct));
    // stepper.f90:113-146 terms
    T c = T(k.cosa), s = T(k.sina);
    o.fcn_c = double(fcn * c);
    o.fcr_s = double(fcr * s);
    o.fcn_s = double(fcn * s);
    o.fcr_c = double(fcr * c);
    o.ri_fct = double(T(fi.r) * fct);
    o.rj_fct = double(T(fj.r) * fct);
    o.mcc = double(mcc);
    double ac = delt_ridge * fmin(fi.h, fj.h);
    o.ac = ac;
    o.fcn_ac = double(fcn) * ac;
}

// Bond of a pair: bond_forces_euler (bonds.f90:74-161) and bond_breaking
// (:164-212) when bonded; reset_bond (reset.f90:18-34) when a record exists
// but the bond is broken.  Strain accumulators and damage stay FP64.
template <bool MIXED>
__device__ __forceinline__ void bond_part(const DevParams& P, const Floe& fi, const Floe& fj,
                                          const Kin& k, PairState& ps, EventOut& o) {
    using T = typename Prec<MIXED>::T;
    if (ps.bond == 1) {
        // every parameter of the bond law is requested here, at once: the
        // tests below are selects, so no parameter load waits in a branch
        const double dt = P.dt, dtd = P.dtd, dth = P.dth, dmax = P.dmax;
        const T eb = T(P.eb), tau_c = T(P.tau_crit), scb_c = T(P.sigmacb_crit);
        const T stb_c = T(P.sigmatb_crit), llb = T(P.lambda_lb);
        ps.deltanb = k.veln * dt + ps.deltanb;
        ps.deltatb = k.veltb * dt + ps.deltatb;
        ps.trb_ji = fi.omega * dt + ps.trb_ji;
        ps.trb_ij = fj.omega * dt + ps.trb_ij;
        T m_redu = T(fi.mass) * T(fj.mass) / (T(fi.mass) + T(fj.mass));
        T i_redu = T(fi.inertia) * T(fj.inertia) / (T(fi.inertia) + T(fj.inertia));
        T L = T(ps.lb);
        T EA = (T(1) - T(ps.damage)) * eb * T(ps.sb);
        T EI = (T(1) - T(ps.damage)) * eb * T(ps.ib);
        T L2 = L * L, L3 = L * L * L;
        T k_axial = EA / L;
        T k_shear1 = T(12) * EI / L3;
        T k_shear2 = T(6) * EI / L2;
        T k_rot1 = T(4) * EI / L;
        T k_rot2 = T(2) * EI / L;
        T eta_i = T(2) * gsqrt(EI * i_redu / L);
        T eta_a = T(2) * gsqrt(EA * m_redu * L);
        T gam_axial = eta_a / L;
        T gam_shear1 = T(12) * eta_i / L3;
        T gam_shear2 = T(6) * eta_i / L2;
        T gam_rot1 = T(4) * eta_i / L;
        T gam_rot2 = T(2) * eta_i / L;
        T oi = T(fi.omega), oj = T(fj.omega);
        T tji = T(ps.trb_ji), tij = T(ps.trb_ij), dtb = T(ps.deltatb), vtb = T(k.veltb);
        T fbn = k_axial * T(ps.deltanb) + gam_axial * T(k.veln);
        T fbt = k_shear1 * dtb + k_shear2 * (tji + tij) + gam_shear1 * vtb + gam_shear2 * (oi + oj);
        T mbb_ji = k_rot1 * tji + k_rot2 * tij - k_shear2 * dtb + gam_rot1 * oi + gam_rot2 * oj
                 - gam_shear2 * vtb;
        T mbb_ij = k_rot2 * tji + k_rot1 * tij - k_shear2 * dtb + gam_rot2 * oi + gam_rot1 * oj
                 - gam_shear2 * vtb;
        // bond_breaking, bonds.f90:164-212
        T sb = T(ps.sb), rb = T(ps.rb), ib = T(ps.ib), hb = T(ps.hb);
        T taub = gabs(fbt) / sb;
        T sigmatb = -fbn / sb + gabs(mbb_ji) * rb / ib;
        T sigmacb = fbn / sb + gabs(mbb_ji) * rb / ib;
        T t = taub / (tau_c * hb);
        T phi = t * t;
        const T qcb = sigmacb / (scb_c * hb), qtb = sigmatb / (stb_c * hb);
        phi = sigmacb > 0 ? phi + qcb * qcb : phi;
        phi = sigmatb > 0 ? phi + qtb * qtb : phi;
        T psi = dmin(T(1), T(1) / gsqrt(phi));
        double D = ps.damage;
        double psid = double(psi);
        ps.damage = D + (1.0 - psid) * (1.0 - D) * dt / dtd - D * dt / dth;
        if (ps.damage >= dmax) {
            ps.bond = 0;
            fbn = T(0); fbt = T(0); mbb_ji = T(0);
            ps.damage = 1.0;
        }
        o.sbd = double(gsqrt(fbn * fbn + fbt * fbt));
        // the terms are written either way (read only where EV_BONDED is set)
        if (ps.bond == 1) o.flags |= EV_BONDED;
        T c = T(k.cosa), s = T(k.sina);
        o.fbn_c = double(fbn * c);
        o.fbt_s = double(fbt * s);
        o.fbn_s = double(fbn * s);
        o.fbt_c = double(fbt * c);
        o.lri_fbt = double(llb * T(fi.r) * fbt);
        o.lrj_fbt = double(llb * T(fj.r) * fbt);
        o.mbb_ji = double(mbb_ji);
        o.mbb_ij = double(mbb_ij);
        o.sb = ps.sb;
        o.fbn_sb = double(fbn) * ps.sb;
    } else if (ps.has_bond_rec) {
        ps.trb_ji = ps.trb_ij = ps.deltanb = ps.deltatb = 0.0;
    }
}

// Sheltering, forcing.f90:162-186 (wind of the outer floe i).  side 1: only
// floe i's coefficients hsf*(j,i); side 2: only floe j's hsf*(i,j); 0: both.
template <bool MIXED>
__device__ __forceinline__ void shelter_part(const DevParams& P, const Floe& fi, const Floe& fj,
                                             const Kin& k, double uai, double vai, double uwi,
                                             double vwi, EventOut& o, int side) {
    using T = typename Prec<MIXED>::T;
    if (!P.shelter) return;
    o.flags |= EV_SHELTER;
    T dn = T(k.deltan), c = T(k.cosa), s = T(k.sina);
    if (side != 1) {
        o.hsfa_ij = double(S_shelter(T(fi.hfa), T(fj.hfa), dn, c, s, T(uai), T(vai)));
        o.hsfw_ij = double(S_shelter(T(fi.hfw), T(fj.hfw), dn, c, s, T(uwi), T(vwi)));
    }
    if (side != 2) {
        o.hsfa_ji = double(S_shelter(T(fj.hfa), T(fi.hfa), dn, -c, -s, T(uai), T(vai)));
        o.hsfw_ji = double(S_shelter(T(fj.hfw), T(fi.hfw), dn, -c, -s, T(uwi), T(vwi)));
    }
}

// |(fcn, fct)| of a pair not in contact: its last contact forces.
template <bool MIXED>
__device__ __forceinline__ double stale_contact_magnitude(const PairState& ps) {
    using T = typename Prec<MIXED>::T;
    const T a = T(ps.fcn), b = T(ps.fct);
    return double(gsqrt(a * a + b * b));
}

// One pair event (i < j), the body of stepper.f90:89-271 for one (j,i).
// fi/fj are updated in place if ridging reshapes one of them.  ps is updated
// to the pair's new state.  Returns the gather terms in `o`.
template <bool MIXED>
__device__ __forceinline__ void pair_event(const DevParams& P, int tstep, int gi, int gj,
                                           Floe& fi, Floe& fj, double uai, double vai,
                                           double uwi, double vwi, PairState& ps,
                                           EventOut& o, int side = 0) {
    o.flags = 0;
    o.ridged = -1;
    o.sbd = 0.0;
    Kin k;
    kinematics(fi, fj, k);
    bond_create(P, tstep, gi, gj, fi, fj, k, ps);
    if (k.deltan > 0) {
        contact_part<MIXED>(P, fi, fj, k, ps, o);
    } else {
        // reset_contact, reset.f90:1-15; the last contact forces stay
        ps.thetarelc = 0.0;
        ps.deltat = 0.0;
        o.sc = stale_contact_magnitude<MIXED>(ps);
    }
    bond_part<MIXED>(P, fi, fj, k, ps, o);
    shelter_part<MIXED>(P, fi, fj, k, uai, vai, uwi, vwi, o, side);
    o.sri = fi.r;
}

// Walls of one floe: verify_bc (boundaries.f90:1-97) -> check_wall (:100-146)
// -> contact_bc (contact_forces.f90:90-244) -> plastic_contact_bc /
// update_shape_bc (ridging.f90:102-179), coulomb_bc (physics.f90:36-70).
// f is reshaped in place if a wall ridges it; hist holds deltat_bc1,
// deltat_bc2, theta_bc1, theta_bc2 and is updated.
template <bool MIXED>
__device__ __forceinline__ void contact_bc(const DevParams& P, Floe& f, int dir1, int dir2, int bd,
                                           double hist[4], double& fn_bc, double& ft_bc,
                                           double& fr_bc, double& mc_bc, double& delt_ridge_bc,
                                           bool& reshaped) {
    using T = typename Prec<MIXED>::T;
    double c = (double)(dir1 * (dir2 - (1 - dir2)));
    double sn = (double)((1 - dir1) * (dir2 - (1 - dir2)));
    double d1 = (double)dir1, d2 = (double)dir2, e1 = (double)(1 - dir1), e2 = (double)(1 - dir2);
    double veln_bc = -f.u * c - f.v * sn;
    double velt_bc = f.u * sn - f.v * c - f.omega * f.r;
    double ri = f.r;
    double deltan_bc = (e2 * (ri - f.x) + d2 * (ri + f.x - P.nx)) * d1
                     + (e2 * (ri - f.y) + d2 * (ri + f.y - P.ny)) * e1;
    double deltat_bc = 0.0;
    double tt = ri - deltan_bc;
    delt_ridge_bc = 2.0 * sqrt(ri * ri - tt * tt);
    if (bd == 1) { hist[0] = -velt_bc * P.dt + hist[0]; deltat_bc = hist[0]; }
    else { hist[1] = -velt_bc * P.dt + hist[1]; deltat_bc = hist[1]; }
    if (bd == 1) hist[2] = -f.omega * P.dt + hist[2];
    else hist[3] = -f.omega * P.dt + hist[3];
    T hi = T(f.h);
    T knc = T(P.pi_ec) * hi * fit(T(deltan_bc) * T(ri) / (T(1.0 * 2.0) * (hi * hi)));
    T ktc = T(P.k6) * knc;
    T dr = T(delt_ridge_bc);
    T dr2 = dr * dr;
    T krc = knc * dr2 / T(12);
    T gamn = -(T(P.beta) * gsqrt(T(4) * knc * T(f.mass) / T(1)));
    T gamt = -(T(P.two_beta) * gsqrt(T(2.0 / 3.0) * ktc * T(f.mass) / T(1)));
    T gamr = gamn * dr2 / T(12);
    T fn = dmax(knc * T(deltan_bc) - gamn * T(veln_bc), T(0));
    T ft = ktc * T(deltat_bc) - gamt * T(velt_bc);
    if (P.ridging) {
        if (T(P.sigmanc_crit) * hi <= fn / dr / hi) {
            // plastic_contact_bc: its dashpots use m(i), the total moment,
            // which reset_forces zeroed for the step (reset.f90:97).
            const T m_total_during_loop = T(0);
            T h2 = hi * hi;
            T kp = T(P.sigmanc_crit) * h2 * dr / T(deltan_bc);
            ktc = T(P.k6) * kp;
            krc = kp * dr2 / T(12);
            gamn = -(T(P.beta) * gsqrt(T(4) * kp * m_total_during_loop / T(1)));
            gamt = -(T(P.two_beta) * gsqrt(T(2.0 / 3.0) * ktc * m_total_during_loop / T(1)));
            gamr = gamn * dr2 / T(12);
            fn = dmin(fn, T(P.sigmanc_crit) * h2 * dr);
            ft = ktc * T(deltat_bc) - gamt * T(velt_bc);
            // update_shape_bc (FP64)
            double drb = delt_ridge_bc;
            double hh = f.h;
            double area = ri * ri * ridge_asin<MIXED>(drb / 2.0 / ri) - drb * (ri - deltan_bc) / 2.0;
            double vol = fmax(area * hh, 0.0);
            double dh = vol / (P.pi * (ri * ri));
            f.r = ri * sqrt(hh / (hh + dh));
            f.h = hh + dh;
            floe_properties(P, f);
            double rn = f.r, tn = rn - deltan_bc;
            delt_ridge_bc = 2.0 * sqrt(rn * rn - tn * tn);
            dr = T(delt_ridge_bc);
            reshaped = true;
        }
    }
    // coulomb_bc
    T fr;
    double* dtb = (bd == 1) ? &hist[0] : &hist[1];
    if (gabs(ft) > T(P.friction_coeff) * gabs(fn)) {
        ft = -(T(P.friction_coeff) * gabs(fn) * gcopysign(T(1), T(velt_bc)));
        fr = ft;
        if (P.slipping) *dtb = double((ft + gamt * T(velt_bc)) / ktc);
    } else {
        fr = ft;
    }
    double th = (bd == 1) ? hist[2] : hist[3];
    T mrolling = krc * T(th) - gamr * T(f.omega);
    if (gabs(T(th)) > T(2) * gabs(fn) / knc / dr) mrolling = T(0);
    fn_bc = double(fn);
    ft_bc = double(ft);
    fr_bc = double(fr);
    mc_bc = double(mrolling);
}

template <bool MIXED>
__device__ __forceinline__ void walls(const DevParams& P, Floe& f, double hist[4], WallOut& w) {
    w.flags = 0;
    w.fx_bc = 0.0; w.fy_bc = 0.0; w.m_bc = 0.0; w.ta_bc = 0.0; w.p_bc = 0.0;
    w.sxx_bc = 0.0; w.syy_bc = 0.0; w.sxy_bc = 0.0; w.syx_bc = 0.0;
    int t1 = 0, t2 = 0;
    const int dirs[4][3] = {{0, 0, 1}, {0, 1, 1}, {1, 0, 2}, {1, 1, 2}};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        bool hit;
        if (k == 0) hit = f.y - f.r < 0;
        else if (k == 1) hit = f.y + f.r > P.ny;
        else if (k == 2) hit = f.x - f.r < 0;
        else hit = f.x + f.r > P.nx;
        if (!hit) continue;
        if (k < 2) t1 = 1; else t2 = 1;
        double fn, ft, fr, mcb, drb;
        bool reshaped = false;
        int d1 = dirs[k][0], d2 = dirs[k][1], bd = dirs[k][2];
        contact_bc<MIXED>(P, f, d1, d2, bd, hist, fn, ft, fr, mcb, drb, reshaped);
        if (reshaped) w.flags |= 4u;
        w.flags |= 1u;
        // check_wall, boundaries.f90:100-146
        double c = (double)(d1 * (d2 - (1 - d2)));
        double sn = (double)((1 - d1) * (d2 - (1 - d2)));
        w.fx_bc = w.fx_bc - fn * c + fr * sn;
        w.fy_bc = w.fy_bc - fn * sn - fr * c;
        w.m_bc = w.m_bc - mcb - f.r * ft;
        // wall stress, boundaries.f90:100-146 (each wall overwrites it)
        const double fnt = sqrt(fn * fn + ft * ft);
        w.sxx_bc = -fnt * f.r * (c * c);
        w.syy_bc = -fnt * f.r * (sn * sn);
        w.sxy_bc = -fnt * f.r * c * sn;
        w.syx_bc = -fnt * f.r * sn * c;
        w.ta_bc = w.ta_bc + drb * f.h;
        w.p_bc = w.p_bc - fn * drb * f.h;
    }
    // reset_boundary calls, boundaries.f90:79-95
    if (t1 + t2 == 0) { hist[0] = hist[1] = hist[2] = hist[3] = 0.0; }
    if (t2 == 0) { hist[3] = 0.0; hist[1] = 0.0; }
    if (t1 == 0) { hist[2] = 0.0; hist[0] = 0.0; }
    w.deltat_bc1 = hist[0]; w.deltat_bc2 = hist[1]; w.theta_bc1 = hist[2]; w.theta_bc2 = hist[3];
    w.nr = f.r; w.nh = f.h; w.nmass = f.mass; w.ninertia = f.inertia; w.nhfa = f.hfa; w.nhfw = f.hfw;
}

// forcing (forcing.f90:32-159) + coriolis (:189-206): per-floe external forces.
struct Forcing {
    double fax, fay, fwx, fwy, ma, mw, fcorx, fcory;
};

template <bool MIXED>
__device__ __forceinline__ void forcing(const DevParams& P, const Floe& f, double uai, double vai,
                                        double uwi, double vwi, double sca_d, double scw_d,
                                        Forcing& o) {
    using T = typename Prec<MIXED>::T;
    // the parameters are requested at once, and the zero-speed and
    // low-speed cases are selects (no parameter load waits in a branch)
    const T z0w = T(P.z0w), l10 = T(P.log10z0), pi = T(P.pi);
    const T rhoair = T(P.rhoair), rhow = T(P.rhowater);
    const T cda = T(P.Cdair), csa = T(P.Csair), cdw = T(P.Cdwater), csw = T(P.Cswater);
    const double sinlat = P.sinlat;
    T sca = T(sca_d), scw = T(scw_d);
    T dua = T(uai - f.u), dva = T(vai - f.v), duw = T(uwi - f.u), dvw = T(vwi - f.v);
    T avn = gsqrt(dua * dua + dva * dva);
    T ovn = gsqrt(duw * duw + dvw * dvw);
    T la = glog(T(f.hfa) / z0w) / l10;
    T lw = glog(T(f.hfw) / z0w) / l10;
    T alog = la * la, olog = lw * lw;
    T r = T(f.r), om = T(f.omega), aom = gabs(om);
    T r2 = r * r;
    const T hfa = T(f.hfa), hfw = T(f.hfw);
    // (a zero relative speed gives zero drag: the terms below are not used)
    const bool a0 = avn == T(0), w0 = ovn == T(0);
    T fdax = rhoair * cda * hfa * r * dua * avn * alog * sca * pi / T(2);
    T fday = rhoair * cda * hfa * r * dva * avn * alog * sca * pi / T(2);
    T fsax = rhoair * pi * r2 * csa * dua / avn * (avn * avn + r2 * (aom * aom) / T(4));
    T fsay = rhoair * pi * r2 * csa * dva / avn * (avn * avn + r2 * (aom * aom) / T(4));
    T fdwx = rhow * cdw * hfw * r * duw * ovn * olog * scw * pi / T(2);
    T fdwy = rhow * cdw * hfw * r * dvw * ovn * olog * scw * pi / T(2);
    T fswx = rhow * pi * r2 * csw * duw / ovn * (ovn * ovn + r2 * (aom * aom) / T(4));
    T fswy = rhow * pi * r2 * csw * dvw / ovn * (ovn * ovn + r2 * (aom * aom) / T(4));
    fdax = a0 ? T(0) : fdax; fday = a0 ? T(0) : fday;
    fsax = a0 ? T(0) : fsax; fsay = a0 ? T(0) : fsay;
    fdwx = w0 ? T(0) : fdwx; fdwy = w0 ? T(0) : fdwy;
    fswx = w0 ? T(0) : fswx; fswy = w0 ? T(0) : fswy;
    o.fax = double(fdax + fsax);
    o.fay = double(fday + fsay);
    o.fwx = double(fdwx + fswx);
    o.fwy = double(fdwy + fswy);
    T r4 = r * r * r * r, r5 = r * r * r * r * r;
    const T ma_lo = -(T(2) * pi / T(5) * r5 * rhoair * csa * om * aom);
    const T ma_hi = -(T(3.0 / 4.0) * pi * rhoair * csa * avn * om * r4);
    const T mw_lo = -(T(2) * pi / T(5) * r5 * rhow * csw * om * aom);
    const T mw_hi = -(T(3.0 / 4.0) * pi * rhow * csw * ovn * om * r4);
    o.ma = double(avn < T(1e-2) ? ma_lo : ma_hi);
    o.mw = double(ovn < T(1e-3) ? mw_lo : mw_hi);
    // coriolis (omega_earth = 7.2921d-5), FP64
    o.fcorx = f.mass * 2.0 * 7.2921e-5 * sinlat * f.v;
    o.fcory = -(f.mass * 2.0 * 7.2921e-5 * sinlat * f.u);
}

// Bilinear sample of one frame; the same expression as the generator's
// sample_grid.
__device__ __forceinline__ void wind_bilinear(const WindGridDesc& g, const double* gu, double x,
                                              double y, double& ua, double& va) {
    double fx = (x - g.x0) / g.dx, fy = (y - g.y0) / g.dy;
    fx = fmin(fmax(fx, 0.0), (double)g.nxn - 1.000001);
    fy = fmin(fmax(fy, 0.0), (double)g.nyn - 1.000001);
    int a = (int)floor(fx), b = (int)floor(fy);
    double tx = fx - a, ty = fy - b;
    const double* gv = gu + (int64_t)g.nxn * g.nyn;
    int64_t k00 = (int64_t)b * g.nxn + a;
    int64_t k01 = k00 + g.nxn;
    double lo = gu[k00] + (gu[k00 + 1] - gu[k00]) * tx;
    double hi = gu[k01] + (gu[k01 + 1] - gu[k01]) * tx;
    ua = lo + (hi - lo) * ty;
    lo = gv[k00] + (gv[k00 + 1] - gv[k00]) * tx;
    hi = gv[k01] + (gv[k01 + 1] - gv[k01]) * tx;
    va = lo + (hi - lo) * ty;
}

// The field at (x, y) at model time t: bilinear in space, linear in time
// between the two frames around t.
__device__ __forceinline__ void wind_point(const WindGridDesc& g, const double* pool, int f0,
                                           double w, double x, double y, double& ua, double& va) {
    const double* gu = pool + g.offset + (int64_t)f0 * 2 * g.nxn * g.nyn;
    wind_bilinear(g, gu, x, y, ua, va);
    if (w > 0.0) {
        double ub, vb;
        wind_bilinear(g, gu + (int64_t)2 * g.nxn * g.nyn, x, y, ub, vb);
        ua = ua + (ub - ua) * w;
        va = va + (vb - va) * w;
    }
}

// The wind a floe of radius r at (x, y) takes at model time t: the field
// at its centre, or (avg) its mean over the floe's disc by the 7-point rule
// of degree 5 for the disc (the centre with weight 1/4, six points at
// radius r sqrt(2/3) and angles k 60 degrees with weight 1/8 each; exact for
// polynomials of degree 5, so for the bilinear field wherever the disc lies
// in one grid cell).  The winds_currents hook of the reference
// (forcing.f90:1-29) plans this averaging over the floe's area.
__device__ __forceinline__ void sample_wind(const WindGridDesc& g, const double* pool, double x,
                                            double y, double r, double t, double& ua, double& va) {
    int f0 = 0;
    double w = 0.0;
    if (g.nframes > 1) {
        const double sf = (t - g.t0) / g.fdt;
        if (sf >= (double)(g.nframes - 1)) {
            f0 = g.nframes - 1;
        } else if (sf > 0.0) {
            f0 = (int)floor(sf);
            w = sf - (double)f0;
        }
    }
    wind_point(g, pool, f0, w, x, y, ua, va);
    if (g.avg && r > 0.0) {
        const double s3 = 0x1.bb67ae8584caap-1;   // sqrt(3)/2
        const double rho = r * 0x1.a20bd700c2c3ep-1;   // r sqrt(2/3)
        const double cs[6][2] = {{1.0, 0.0}, {0.5, s3}, {-0.5, s3}, {-1.0, 0.0}, {-0.5, -s3}, {0.5, -s3}};
        double su = 0.25 * ua, sv = 0.25 * va;
        for (int q = 0; q < 6; ++q) {
            double pu, pv;
            wind_point(g, pool, f0, w, x + rho * cs[q][0], y + rho * cs[q][1], pu, pv);
            su = su + 0.125 * pu;
            sv = sv + 0.125 * pv;
        }
        ua = su;
        va = sv;
    }
}

}  // namespace godar
