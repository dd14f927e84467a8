// Correctly rounded double exp, log, asin and acos (round to nearest) for
// the FP64 force laws.  The reference's transcendental calls
// (forcing.f90:222 log, forcing.f90:261 exp, ridging.f90:56-57 acos,
// ridging.f90:155 asin) go to the host's libm; a device libm that rounds
// differently in a few percent of calls lets the chaotic trajectories part
// after some thousand steps.  These functions return the double nearest to
// the exact value, so they differ from a libm only where that libm itself
// misrounds (about 0.03-0.2 % of calls for glibc's), which leaves the
// trajectories equal far longer.
//
// Method: evaluation in double-double arithmetic (hi + lo pairs with
// error-free sums and products, the products through fma; Dekker 1971,
// Knuth TAOCP vol. 2 4.2.2) to about 100 bits, then one rounding of hi + lo.
// The result is the correctly rounded value unless the exact value lies
// within ~2^-100 (relative) of a rounding boundary.
//   exp:  x = (64 k + j) ln2/64 + r, |r| <= ln2/128 (Cody-Waite with a
//         three-part ln2/64), exp(x) = 2^k 2^(j/64) exp(r), exp(r) by its
//         Taylor series (terms above r^5 in double, the rest in
//         double-double).
//   log:  one Newton step from the device log y0: log x = y0 + log1p(t),
//         t = x exp(-y0) - 1 with exp in double-double; for x near 1 the
//         log1p series of x - 1 directly.
//   asin: one Newton step from the device asin w0 on sin(w) = z with sin and
//         cos of w0 in double-double (Taylor series of w0/4 and two
//         double-angle steps); |z| > 1/2 through asin z = pi/2 - 2 asin
//         sqrt((1 - z)/2).  acos from asin: acos x = pi/2 - asin x,
//         2 asin sqrt((1 - x)/2) or pi - 2 asin sqrt((1 + x)/2).
// exp results below 2^-1022 (subnormal) are rounded twice; the force laws
// only use such values in 1 - exp(x), where they vanish.
#pragma once
#include <cmath>
#include <cstdint>

#include "cr_tables.h"

#if defined(__HIPCC__)
#define CRM_HD __host__ __device__ __forceinline__
#else
#define CRM_HD inline
#endif

namespace crm {

struct dd {
    double hi, lo;
};

CRM_HD dd two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
// |a| >= |b| (or a == 0)
CRM_HD dd fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
CRM_HD dd two_prod(double a, double b) {
    const double p = a * b;
    return {p, std::fma(a, b, -p)};
}
CRM_HD dd dd_add(dd a, dd b) {
    dd s = two_sum(a.hi, b.hi);
    const dd t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return fast_two_sum(s.hi, s.lo);
}
CRM_HD dd dd_add_d(dd a, double b) {
    dd s = two_sum(a.hi, b);
    s.lo += a.lo;
    return fast_two_sum(s.hi, s.lo);
}
CRM_HD dd dd_mul(dd a, dd b) {
    dd p = two_prod(a.hi, b.hi);
    p.lo += a.hi * b.lo;
    p.lo += a.lo * b.hi;
    return fast_two_sum(p.hi, p.lo);
}
CRM_HD dd dd_mul_d(dd a, double b) {
    dd p = two_prod(a.hi, b);
    p.lo += a.lo * b;
    return fast_two_sum(p.hi, p.lo);
}
CRM_HD dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
CRM_HD dd fact_inv(int k) { return {kInvFact[k][0], kInvFact[k][1]}; }

// exp(x) = 2^k * e (e a double-double in [0.7, 1.5)), for finite
// x in [-746, 710]
CRM_HD dd exp_dd(double x, int& k) {
    const double n = std::rint(x * kExpInvL);
    const int ni = (int)n;
    const int j = ni & 63;
    k = (ni - j) / 64;
    // r = x - n ln2/64: n*L1 is exact and cancels exactly against x
    const double t1 = x - n * kExpL1;
    const dd p2 = two_prod(n, kExpL2);
    dd r = two_sum(t1, -p2.hi);
    r.lo -= p2.lo;
    r.lo -= n * kExpL3;
    r = fast_two_sum(r.hi, r.lo);
    // exp(r) = sum r^i / i!, i <= 11: i >= 6 in double (each term < 2^-53
    // of the sum), i <= 5 by a double-double Horner scheme
    const double rh = r.hi;
    double tail = kInvFact[11][0];
#pragma unroll
    for (int i = 10; i >= 6; --i) tail = kInvFact[i][0] + rh * tail;
    dd p = dd_mul_d(r, tail);
#pragma unroll
    for (int i = 5; i >= 1; --i) p = dd_mul(r, dd_add(fact_inv(i), p));
    p = dd_add_d(p, 1.0);
    return dd_mul({kExp2Tab[j][0], kExp2Tab[j][1]}, p);
}

CRM_HD double cr_exp(double x) {
    if (x != x) return x + x;
    if (x > 0x1.62e42fefa39efp+9) return HUGE_VAL;   // above log(DBL_MAX)
    if (x < -0x1.74910d52d3052p+9) return 0.0;       // below log(2^-1075)
    if (std::fabs(x) < 0x1.0p-26) {
        // 1 + x + tau, tau = x^2/2 + x^3/6 (the rest is below 2^-104): 1 + x
        // is exact as t.hi + t.lo, and t.hi is the answer unless t.lo + tau
        // reaches half an ulp above it (tau > 0 and t.lo >= -half an ulp
        // below, so it never rounds down); the differences are exact
        const dd t = two_sum(1.0, x);
        const double tau = 0.5 * x * x + x * x * x / 6.0;
        const double up = std::nextafter(t.hi, HUGE_VAL);
        return tau > (up - t.hi) * 0.5 - t.lo ? up : t.hi;
    }
    int k;
    const dd e = exp_dd(x, k);
    // e is normalised: e.hi is the double nearest to e.hi + e.lo
    return std::ldexp(e.hi, k);
}

// log(1 + u) for |u| < 2^-6 by its series, double-double: sum over
// i = 1..19 of (-1)^(i+1) u^i / i (u^20/20 < 2^-126 |u|); the terms above u^8
// (each below 2^-47 of the sum) in double, the rest by a double-double
// Horner scheme with the coefficients 1/i to double-double precision
CRM_HD dd log1p_small(double u) {
    double tail = 0.0;
#pragma unroll
    for (int i = 19; i >= 9; --i) tail = ((i & 1) ? 1.0 : -1.0) / i + u * tail;
    dd p = two_prod(u, tail);
#pragma unroll
    for (int i = 8; i >= 1; --i) {
        const double hi = 1.0 / i;
        const dd ci = {hi, std::fma(-hi, (double)i, 1.0) / i};   // 1/i = hi + (1 - hi i)/i
        p = dd_mul_d(dd_add((i & 1) ? ci : dd_neg(ci), p), u);
    }
    return p;
}

CRM_HD double cr_log(double x) {
    if (x != x) return x + x;
    if (x < 0.0) return __builtin_nan("");
    if (x == 0.0) return -HUGE_VAL;
    if (x == HUGE_VAL) return x;
    if (std::fabs(x - 1.0) < 0x1.0p-6) {
        const dd y = log1p_small(x - 1.0);    // x - 1 exact (Sterbenz)
        return y.hi;
    }
    double s = x;
    double shift = 0.0;
    if (x < 0x1.0p-1000) {                      // keep x 2^-k normal below
        s = x * 0x1.0p+200;
        shift = -200.0;
    }
    const double y0 = std::log(s);
    int k;
    const dd e = exp_dd(-y0, k);                // exp(-y0) = 2^k e
    // t = s exp(-y0) - 1, small: s 2^k is exact
    const double sk = std::ldexp(s, k);
    dd t = two_prod(sk, e.hi);
    t.lo += sk * e.lo;
    t = dd_add_d(t, -1.0);
    // log s = y0 + t - t^2/2 (t^3/3 < 2^-140)
    dd y = two_sum(y0, t.hi);
    y.lo += t.lo - 0.5 * t.hi * t.hi;
    y = fast_two_sum(y.hi, y.lo);
    if (shift != 0.0) {
        // + shift ln2 (rare: x below 2^-1000)
        const dd ln2 = {0x1.62e42fefa39efp-1, 0x1.abc9e3b39803fp-56};
        y = dd_add(y, dd_mul_d(ln2, shift));
    }
    return y.hi;
}

// sin and cos of w, |w| <= 0.8, double-double (w / 4 by its Taylor series,
// then two double-angle steps)
CRM_HD void sincos_dd(dd w, dd& s, dd& c) {
    const dd a = {w.hi * 0.25, w.lo * 0.25};
    const dd a2 = dd_mul(a, a);
    // sin a = a (1 - a^2/3! + a^4/5! - ...), cos a = 1 - a^2/2! + a^4/4! - ...
    // to a^23 / a^22; terms from a^12 on in double
    const double a2h = a2.hi;
    double ts = 0.0, tc = 0.0;
#pragma unroll
    for (int i = 11; i >= 6; --i) {
        const double sg = (i & 1) ? -1.0 : 1.0;
        ts = sg * kInvFact[2 * i + 1][0] + a2h * ts;
        tc = sg * kInvFact[2 * i][0] + a2h * tc;
    }
    dd ps = dd_mul_d(a2, ts), pc = dd_mul_d(a2, tc);
#pragma unroll
    for (int i = 5; i >= 1; --i) {
        const dd fs = fact_inv(2 * i + 1), fc = fact_inv(2 * i);
        ps = dd_mul(a2, dd_add((i & 1) ? dd_neg(fs) : fs, ps));
        pc = dd_mul(a2, dd_add((i & 1) ? dd_neg(fc) : fc, pc));
    }
    dd sa = dd_mul(a, dd_add_d(ps, 1.0));
    dd ca = dd_add_d(pc, 1.0);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const dd s2 = dd_mul(sa, ca);
        const dd c2 = dd_add(dd_mul(ca, ca), dd_neg(dd_mul(sa, sa)));
        sa = {2.0 * s2.hi, 2.0 * s2.lo};
        ca = c2;
    }
    s = sa;
    c = ca;
}

// asin z for |z| <= 0.7072 (z a double-double), double-double
CRM_HD dd asin_small(dd z) {
    const double w0 = std::asin(z.hi);
    dd s, c;
    sincos_dd({w0, 0.0}, s, c);
    // w = w0 + (z - sin w0) / cos w0 (the next term, tan(w) d^2 / 2, is below 2^-104)
    const dd num = dd_add(z, dd_neg(s));
    const double d = (num.hi + num.lo) / c.hi;
    dd w = two_sum(w0, d);
    return fast_two_sum(w.hi, w.lo);
}

// sqrt((1 - |x|) / 2) for 1/2 <= |x| <= 1, double-double
CRM_HD dd half_versine_sqrt(double ax) {
    const double h = (1.0 - ax) * 0.5;   // exact for ax in [1/2, 1]
    const double q = std::sqrt(h);
    if (q == 0.0) return {0.0, 0.0};
    const double e = std::fma(-q, q, h);  // h - q^2, exact
    return fast_two_sum(q, e / (2.0 * q));
}

CRM_HD double cr_asin(double x) {
    if (x != x) return x + x;
    const double ax = std::fabs(x);
    if (ax > 1.0) return __builtin_nan("");
    if (ax <= 0.5) return asin_small({x, 0.0}).hi;
    // asin |x| = pi/2 - 2 asin sqrt((1 - |x|)/2)
    const dd w = asin_small(half_versine_sqrt(ax));
    const dd r = dd_add({kPio2Hi, kPio2Lo}, {-2.0 * w.hi, -2.0 * w.lo});
    return x < 0.0 ? -r.hi : r.hi;
}

CRM_HD double cr_acos(double x) {
    if (x != x) return x + x;
    const double ax = std::fabs(x);
    if (ax > 1.0) return __builtin_nan("");
    if (ax <= 0.5) {
        const dd a = asin_small({x, 0.0});
        return dd_add({kPio2Hi, kPio2Lo}, dd_neg(a)).hi;
    }
    const dd w = asin_small(half_versine_sqrt(ax));
    if (x > 0.0) return fast_two_sum(2.0 * w.hi, 2.0 * w.lo).hi;
    return dd_add({kPiHi, kPiLo}, {-2.0 * w.hi, -2.0 * w.lo}).hi;
}

}  // namespace crm
