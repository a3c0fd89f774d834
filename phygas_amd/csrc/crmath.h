// Correctly rounded FP64 pow, x^1.5, hypot, sin and cos for the device side of the plain FP64 path.
//
// The plain path must reproduce the reference's CPU orbits, whose forces and start states come from the
// host C library (glibc: pow in the Allen & Santillan force, reference src/potential.c:57,71; sin and cos
// in the coordinate transform, src/coords.c:83-85; hypot in the rotation prior, src/phygas.c:26).  glibc
// rounds these to nearest in all but rare cases (its pow is documented at 0.52 ulp).  The device library's
// versions differ in the last bit on a sizeable share of arguments, and the finite-difference
// state-transition matrix (A = (x_j - x_0) / 1e-5) amplifies a one-ulp difference of an 8-kpc position by
// 1e5.  Functions that are correctly rounded agree with glibc wherever glibc is, so the device then
// integrates the same doubles as the host.
//
// Method: double-double arithmetic (Dekker's exact product through fma, Knuth's two-sum) around
// table-driven reductions in the style of P. T. P. Tang (ACM TOMS 15 (1989) 144-157 for exp, 16 (1990)
// 378-400 for log):
//   log x   x = 2^e m, m in [1, 2); c = 1/m_i from a 128-entry table (m_i the centre of m's bucket), so
//           r = m c - 1 is exact as a double-double and |r| <= 2^-8; log x = e ln2 - log c + log1p(r), with
//           log1p through r^8 (truncation below 2^-75) and every term above 2^-70 kept in double-double
//   exp z   z = (64 k' + j) ln2/64 + r, |r| <= ln2/128; exp z = 2^k' 2^(j/64) exp(r), exp(r) through r^8
//           (truncation below 2^-75), 2^(j/64) from a double-double table
//   pow     exp(y log x) with y log x as a double-double: absolute error about 2^-72, so the result's
//           relative error before the final rounding is about 2^-72 and it misrounds only when the exact
//           value lies within that distance of a midpoint between two doubles
//   x^1.5   x sqrt(x) in double-double (sqrt's residual through fma): error below 2^-100
//   hypot   sqrt(x^2 + y^2) with the sum exact in double-double and one correction step of the square root
//   sin/cos x = k pi/2 + r (Cody-Waite with pi/2 in three 33-bit pieces), r = j/64 + t with |t| <= 1/128,
//           sin/cos(j/64) from a double-double table and short series for t
// The final step rounds a double-double once, so each result is the double nearest the true value unless
// that value lies within about 2^-70 (relative) of a midpoint.
//
// Domain (checked; outside it the device library's function is used): pow x in (2^-1000, 2^1000) with
// |y log x| < 700; x^1.5 x in (2^-600, 2^600); hypot |x|, |y| below 2^500 and not both below 2^-500;
// sin/cos |x| < 2^19.  The plain path's arguments lie well inside.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "crmath_tables.h"

#define CR_HD __host__ __device__ __forceinline__

namespace phx {
namespace cr {

struct DD { double hi, lo; };

// a + b = s + e exactly (Knuth)
CR_HD DD two_sum(double a, double b) {
    const double s = a + b, bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
// a + b = s + e exactly, for |a| >= |b| or a == 0 (Dekker)
CR_HD DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
// a * b = p + e exactly (one fma)
CR_HD DD two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}

CR_HD uint64_t bits(double x) { return __builtin_bit_cast(uint64_t, x); }
CR_HD double from_bits(uint64_t b) { return __builtin_bit_cast(double, b); }

// log x as a double-double, x a positive normal double; absolute error below 2^-74.
CR_HD DD log_dd(double x) {
    const uint64_t b = bits(x);
    const int e = (int)(b >> 52) - 1023;
    const double m = from_bits((b & 0x000FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL);   // [1, 2)
    const int i = (int)((b >> 45) & 127);
    const double c = LOG_TAB[i][0], lch = LOG_TAB[i][1], lcl = LOG_TAB[i][2];
    const DD p = two_prod(m, c);
    const double rh = p.hi - 1.0, rl = p.lo;   // exact: p.hi lies in [1 - 2^-8, 1 + 2^-8]
    // log1p(r) = r - r^2/2 + r^3 (1/3 - r/4 + r^2/5 - r^3/6 + r^4/7 - r^5/8)
    const DD sq = two_prod(rh, rh);
    const double poly = 1.0 / 3 - rh * (0.25 - rh * (0.2 - rh * (1.0 / 6 - rh * (1.0 / 7 - rh * 0.125))));
    const double tail = rh * sq.hi * poly;
    const DD a = fast_two_sum(rh, -0.5 * sq.hi);
    const double lo1 = a.lo + rl - 0.5 * sq.lo - rh * rl + tail;
    // e ln2 + (-log c) + log1p(r); e * LN2_HI is exact
    const double ef = (double)e;
    const DD s1 = two_sum(ef * LN2_HI, lch);
    const DD s2 = two_sum(s1.hi, a.hi);
    const double low = s1.lo + s2.lo + lo1 + lcl + ef * LN2_LO;
    return fast_two_sum(s2.hi, low);
}

// exp(zh + zl) rounded once to double; |zh| < 700, |zl| <= 2^-40 |zh| + 2^-60.
CR_HD double exp_dd(double zh, double zl) {
    const double kf = rint(zh * INV_LN2_64);
    const int k = (int)kf;
    const double t = fma(-kf, LN2_64_HI, zh);   // exact: k * LN2_64_HI is exact and within a factor 2 of zh
    const DD r0 = two_sum(t, -kf * LN2_64_LO);
    const DD r = two_sum(r0.hi, r0.lo + zl);
    const double rh = r.hi, rl = r.lo;
    // exp(r) - 1 = rh + rh^2/2 + rh^3 Q(rh) + rl (1 + rh), Q through rh^5 / 8!
    const DD sq = two_prod(rh, rh);
    const double q = 1.0 / 6 + rh * (1.0 / 24 + rh * (1.0 / 120 + rh * (1.0 / 720 + rh * (1.0 / 5040 + rh * (1.0 / 40320)))));
    const DD p = fast_two_sum(rh, 0.5 * sq.hi);
    const double pl = p.lo + 0.5 * sq.lo + rh * sq.hi * q + rl + rl * rh;
    // 2^(j/64) (1 + p)
    const int j = k & 63;
    const double th = EXP_TAB[j][0], tl = EXP_TAB[j][1];
    const DD u = two_prod(th, p.hi);
    const DD s = fast_two_sum(th, u.hi);
    const double v = s.hi + (s.lo + (u.lo + th * pl + tl + tl * p.hi));
    return ldexp(v, k >> 6);   // k >> 6 = floor(k / 64); exact for normal results
}

// x^y, correctly rounded (see above)
CR_HD double pow_cr(double x, double y) {
    if (!(x > 0x1p-1000 && x < 0x1p1000)) return pow(x, y);
    const DD l = log_dd(x);
    const DD z = two_prod(y, l.hi);
    const double zl = z.lo + y * l.lo;
    if (!(fabs(z.hi) < 700.0)) return pow(x, y);
    return exp_dd(z.hi, zl);
}

// x^y1 and x^y2 sharing one logarithm (the halo's two powers)
CR_HD void pow2_cr(double x, double y1, double y2, double &p1, double &p2) {
    if (!(x > 0x1p-1000 && x < 0x1p1000)) {
        p1 = pow(x, y1);
        p2 = pow(x, y2);
        return;
    }
    const DD l = log_dd(x);
    const DD z1 = two_prod(y1, l.hi), z2 = two_prod(y2, l.hi);
    p1 = fabs(z1.hi) < 700.0 ? exp_dd(z1.hi, z1.lo + y1 * l.lo) : pow(x, y1);
    p2 = fabs(z2.hi) < 700.0 ? exp_dd(z2.hi, z2.lo + y2 * l.lo) : pow(x, y2);
}

// x^1.5 = x sqrt(x), correctly rounded
CR_HD double pow15_cr(double x) {
    if (!(x > 0x1p-600 && x < 0x1p600)) return pow(x, 1.5);
    const double s = sqrt(x);
    const double sl = fma(-s, s, x) / (2.0 * s);   // sqrt(x) = s + sl + O(2^-106 s)
    const DD p = two_prod(x, s);
    return p.hi + (p.lo + x * sl);
}

// sqrt(x^2 + y^2), correctly rounded
CR_HD double hypot_cr(double x, double y) {
    const double ax = fabs(x), ay = fabs(y);
    const double big = ax > ay ? ax : ay;
    if (!(big < 0x1p500 && big > 0x1p-500)) return hypot(x, y);
    const DD sx = two_prod(ax, ax), sy = two_prod(ay, ay);
    const DD s0 = two_sum(sx.hi, sy.hi);
    const DD s = fast_two_sum(s0.hi, s0.lo + sx.lo + sy.lo);
    const double r = sqrt(s.hi);
    return r + (fma(-r, r, s.hi) + s.lo) / (2.0 * r);
}

// sin x and cos x, correctly rounded
CR_HD void sincos_cr(double x, double &sn, double &cs) {
    if (!(fabs(x) < 0x1p19)) {
        sn = sin(x);
        cs = cos(x);
        return;
    }
    const double kf = rint(x * TWO_OVER_PI);
    const int k = (int)kf;
    const double t1 = fma(-kf, PIO2_1, x);   // exact
    const DD r0 = two_sum(t1, -kf * PIO2_2);   // kf * PIO2_2 exact
    DD r = two_sum(r0.hi, r0.lo - kf * PIO2_3);
    const bool neg = r.hi < 0;
    if (neg) { r.hi = -r.hi; r.lo = -r.lo; }
    const int j = (int)rint(r.hi * 64.0);   // 0 .. 51
    const double th = r.hi - j * (1.0 / 64), tl = r.lo;   // exact; |t| <= 1/128
    const double t2 = th * th;
    // sin t = t + t^3 (-1/6 + t^2 (1/120 + t^2 (-1/5040 + t^2 / 362880)))
    const double ts = th * t2 * (-1.0 / 6 + t2 * (1.0 / 120 + t2 * (-1.0 / 5040 + t2 * (1.0 / 362880))));
    const DD sv = fast_two_sum(th, ts);
    const double sh = sv.hi, sl = sv.lo + tl;
    // cos t = 1 - t^2/2 + t^4 (1/24 + t^2 (-1/720 + t^2 / 40320))
    const DD sq = two_prod(th, th);
    const double tc = t2 * t2 * (1.0 / 24 + t2 * (-1.0 / 720 + t2 * (1.0 / 40320)));
    const DD cv = fast_two_sum(1.0, -0.5 * sq.hi);
    const double ch = cv.hi, cl = cv.lo - 0.5 * sq.lo - th * tl + tc;
    const double SH = SC_TAB[j][0], SL = SC_TAB[j][1], CH = SC_TAB[j][2], CL = SC_TAB[j][3];
    // sin(a + t) = sin a cos t + cos a sin t, cos(a + t) = cos a cos t - sin a sin t
    const DD u1 = two_prod(SH, ch), u2 = two_prod(CH, sh);
    const DD us = two_sum(u1.hi, u2.hi);
    double s = us.hi + (us.lo + u1.lo + u2.lo + SH * cl + SL * ch + CH * sl + CL * sh);
    const DD w1 = two_prod(CH, ch), w2 = two_prod(SH, sh);
    const DD wc = two_sum(w1.hi, -w2.hi);
    const double c = wc.hi + (wc.lo + w1.lo - w2.lo + CH * cl + CL * ch - SH * sl - SL * sh);
    if (neg) s = -s;
    switch (k & 3) {
    case 0: sn = s; cs = c; break;
    case 1: sn = c; cs = -s; break;
    case 2: sn = -s; cs = -c; break;
    default: sn = -c; cs = s; break;
    }
}

}  // namespace cr
}  // namespace phx
