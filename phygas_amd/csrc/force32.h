// FP32 orbit arithmetic of the fast form, shared by the gfx950 kernels and host-side checks.
//
// * force32():      Allen & Santillan acceleration in FP32 (x^1.5 as x*sqrt(x), the halo powers through
//                   exp2/log2), evaluated at the high part of a float-float position.
// * dforce32():     the DIFFERENCE a(p0 + d) - a(p0) of two nearby positions, written so that every FP32
//                   rounding error is relative to the difference itself (no subtraction of two ~1e-4 forces):
//                   squared radii move by 2 p.d + d.d, and powers move by (1+t)^e - 1 summed as a series in
//                   t = (change of the base) / base, |t| < ~1e-3 on these orbits.  Propagating the six
//                   neighbour offsets with it reproduces the finite-difference state-transition matrix of the
//                   reference (reference: src/phygas.c:286-323: six orbits 1e-5 apart, differenced) to FP32
//                   relative precision of the offsets, not the tangent-linear approximation (which differs
//                   from the finite difference by ~1e-4 relative after 100 Myr).
// * ff_add():       float-float (Kahan-compensated) accumulation, for host measurements of an FP32 nominal orbit
//                   with compensated position and velocity (the kernels keep the nominal in FP64).
//
// Potential: reference src/potential.c:54-101.
#pragma once

#include <hip/hip_runtime.h>

#include "physics.h"

#define PHX_HDF __host__ __device__ __forceinline__

namespace phx {

#if defined(__HIP_DEVICE_COMPILE__)
PHX_HDF float f_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
PHX_HDF float f_sqrt(float x) { return __builtin_sqrtf(x); }
PHX_HDF float f_log2(float x) { return __builtin_amdgcn_logf(x); }
PHX_HDF float f_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
PHX_HDF float f_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
PHX_HDF float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
PHX_HDF float f_sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }   // the instruction alone (~1 ulp)
#else
PHX_HDF float f_rsqrt(float x) { return 1.0f / sqrtf(x); }
PHX_HDF float f_sqrt(float x) { return sqrtf(x); }
PHX_HDF float f_log2(float x) { return log2f(x); }
PHX_HDF float f_exp2(float x) { return exp2f(x); }
PHX_HDF float f_fma(float a, float b, float c) { return fmaf(a, b, c); }
PHX_HDF float f_rcp(float x) { return 1.0f / x; }
PHX_HDF float f_sqrt_hw(float x) { return sqrtf(x); }
#endif
// a / b through the hardware reciprocal (within ~2 ulp; a reciprocal shared by several quotients is
// evaluated once)
PHX_HDF float f_div(float a, float b) { return a * f_rcp(b); }

struct PotF {
    float bulge_gm, b2, disk_gm, disk_a, db2, halo_k1, halo_k2, halo_cut2, halo_gmcut, halo_p1, halo_p2;
    // halo inside the cutoff: f(s) = k1 s^p1 / (1 + k2 s^p2) with s = r^2, p1 = (g-3)/2, p2 = (g-1)/2,
    // k1 = GM_h / a_h^g, k2 = a_h^-(g-1); a_halo = -f(s) p.  Outside: f = GM_enc(cut) / r^3.
};

inline PotF make_potf(const Pot &P) {
    PotF f;
    f.bulge_gm = (float)P.bulge_gm;
    f.b2 = (float)(P.bulge_b * P.bulge_b);
    f.disk_gm = (float)P.disk_gm;
    f.disk_a = (float)P.disk_a;
    f.db2 = (float)(P.disk_b * P.disk_b);
    f.halo_k1 = (float)(P.halo_gm / pow(P.halo_a, P.halo_gamma));
    f.halo_k2 = (float)pow(P.halo_a, -(P.halo_gamma - 1.0));
    f.halo_cut2 = (float)(P.halo_cut * P.halo_cut);
    double xc = P.halo_cut / P.halo_a;
    f.halo_gmcut = (float)(P.halo_gm * pow(xc, P.halo_gamma) / (1.0 + pow(xc, P.halo_gamma - 1.0)));
    f.halo_p1 = (float)((P.halo_gamma - 3.0) / 2.0);
    f.halo_p2 = (float)((P.halo_gamma - 1.0) / 2.0);
    return f;
}

struct F3 { float x, y, z; };

// s^e for s > 0
PHX_HDF float f_pow(float s, float e) { return f_exp2(e * f_log2(s)); }

// Acceleration at p (FP32), its quotients through the hardware reciprocal (within ~2 ulp).
PHX_HDF F3 force32(const PotF &Q, float x, float y, float z) {
    const float R2 = f_fma(x, x, y * y), r2 = f_fma(z, z, R2);
    const float u = r2 + Q.b2;
    const float iu = f_rsqrt(u);
    const float fb = -Q.bulge_gm * iu * iu * iu;
    const float zeta = f_sqrt_hw(f_fma(z, z, Q.db2));
    const float az = Q.disk_a + zeta;
    const float D2 = f_fma(az, az, R2);
    const float iD = f_rsqrt(D2);
    const float iD3 = iD * iD * iD;
    const float fd = -Q.disk_gm * iD3;
    float fh;
    if (r2 < Q.halo_cut2) {
        fh = f_div(-Q.halo_k1 * f_pow(r2, Q.halo_p1), 1.0f + Q.halo_k2 * f_pow(r2, Q.halo_p2));
    } else {
        const float ir = f_rsqrt(r2);
        fh = -Q.halo_gmcut * ir * ir * ir;
    }
    F3 a;
    a.x = (fb + fd + fh) * x;
    a.y = (fb + fd + fh) * y;
    a.z = (fb + fh) * z + fd * z * f_div(az, zeta);
    return a;
}

// (1 + t)^e - 1 for |t| < 1e-2 as a series, relative accuracy ~ FP32 of the result.
PHX_HDF float pw1_series(float t, float e) {
    const float c2 = e * (e - 1.0f) * 0.5f, c3 = c2 * (e - 2.0f) * (1.0f / 3.0f), c4 = c3 * (e - 3.0f) * 0.25f,
                c5 = c4 * (e - 4.0f) * 0.2f;
    return t * f_fma(t, f_fma(t, f_fma(t, f_fma(t, c5, c4), c3), c2), e);
}

// (1 + t)^e - 1: the series for |t| < 1e-2; outside that range the direct form (loses relative precision,
// only reached far from any orbit this path integrates).
PHX_HDF float pw1(float t, float e) {
    if (fabsf(t) < 1e-2f) return pw1_series(t, e);
    return f_pow(1.0f + t, e) - 1.0f;
}

// a(p + d) - a(p), every rounding relative to the difference.  Fast = true takes the common case without
// branches: every (1 + t)^e - 1 through its series and the halo inside its cutoff at both positions; it
// raises tmax to the largest |t| of its series and r2max to the larger squared radius.  The caller checks
// tmax < 1e-2 and r2max < halo_cut2 (a NaN fails the check) and otherwise takes the Fast = false form, which
// branches per term.  Where the check holds both forms run the same operations, so their results are
// identical.
template <bool Fast>
PHX_HDF F3 dforce32_t(const PotF &Q, float x, float y, float z, float dx, float dy, float dz, float &tmax,
                      float &r2max) {
    auto pw = [&](float t, float e) {
        if (Fast) {
            tmax = fmaxf(tmax, fabsf(t));
            return pw1_series(t, e);
        }
        return pw1(t, e);
    };
    const float R2 = f_fma(x, x, y * y), r2 = f_fma(z, z, R2);
    const float dR2 = f_fma(dx, 2.0f * x + dx, dy * (2.0f * y + dy));
    const float dr2 = f_fma(dz, 2.0f * z + dz, dR2);
    // bulge: -GM p u^-3/2, u = r^2 + b^2
    const float u0 = r2 + Q.b2;
    const float iu0 = f_rsqrt(u0), g0 = iu0 * iu0 * iu0;
    const float dg = g0 * pw(f_div(dr2, u0), -1.5f);
    const float g1 = g0 + dg;
    // disk: -GM (x, y, z (a + zeta) / zeta) D^-3, D^2 = R^2 + (a + zeta)^2, zeta = sqrt(z^2 + b^2)
    const float zeta0 = f_sqrt_hw(f_fma(z, z, Q.db2)), zeta1 = f_sqrt_hw(f_fma(z + dz, z + dz, Q.db2));
    const float dzeta = f_div(dz * (2.0f * z + dz), zeta0 + zeta1);
    const float A0 = Q.disk_a + zeta0;
    const float D20 = f_fma(A0, A0, R2);
    const float dD2 = f_fma(dzeta, 2.0f * A0 + dzeta, dR2);
    const float iD0 = f_rsqrt(D20), h0 = iD0 * iD0 * iD0;
    const float dh = h0 * pw(f_div(dD2, D20), -1.5f);
    const float h1 = h0 + dh;
    // z (1 + a / zeta): difference of z / zeta is (dz zeta0 - z dzeta) / (zeta0 zeta1)
    const float q0 = z * (1.0f + f_div(Q.disk_a, zeta0));
    const float dq = dz + f_div(Q.disk_a * (dz * zeta0 - z * dzeta), zeta0 * zeta1);
    // halo
    float f0, df;
    const float r21 = r2 + dr2;
    if (Fast) r2max = fmaxf(r2max, fmaxf(r2, r21));
    if (Fast || (r2 < Q.halo_cut2 && r21 < Q.halo_cut2)) {
        const float t = f_div(dr2, r2);
        const float N0 = f_pow(r2, Q.halo_p1), E0 = Q.halo_k2 * f_pow(r2, Q.halo_p2);
        const float dN = N0 * pw(t, Q.halo_p1), dE = E0 * pw(t, Q.halo_p2);
        const float Dn0 = 1.0f + E0, Dn1 = Dn0 + dE;
        f0 = f_div(Q.halo_k1 * N0, Dn0);
        df = f_div(Q.halo_k1 * (dN * Dn0 - N0 * dE), Dn0 * Dn1);
    } else if (r2 >= Q.halo_cut2 && r21 >= Q.halo_cut2) {
        const float ir = f_rsqrt(r2);
        f0 = Q.halo_gmcut * ir * ir * ir;
        df = f0 * pw(f_div(dr2, r2), -1.5f);
    } else {   // straddles the cutoff: plain difference
        const float ir0 = f_rsqrt(r2), ir1 = f_rsqrt(r21);
        f0 = r2 < Q.halo_cut2 ? Q.halo_k1 * f_pow(r2, Q.halo_p1) / (1.0f + Q.halo_k2 * f_pow(r2, Q.halo_p2))
                              : Q.halo_gmcut * ir0 * ir0 * ir0;
        const float f1 = r21 < Q.halo_cut2
                             ? Q.halo_k1 * f_pow(r21, Q.halo_p1) / (1.0f + Q.halo_k2 * f_pow(r21, Q.halo_p2))
                             : Q.halo_gmcut * ir1 * ir1 * ir1;
        df = f1 - f0;
    }
    const float f1 = f0 + df;
    // a = -(GMb g + GMd h + f) p (x, y) ; a_z = -(GMb g + f) z - GMd h q
    const float kxy0 = Q.bulge_gm * g0 + Q.disk_gm * h0 + f0;
    const float dkxy = Q.bulge_gm * dg + Q.disk_gm * dh + df;
    const float kxy1 = kxy0 + dkxy;
    const float kz0 = Q.bulge_gm * g0 + f0, dkz = Q.bulge_gm * dg + df, kz1 = kz0 + dkz;
    (void)f1; (void)g1; (void)h1;
    F3 d;
    d.x = -(f_fma(dkxy, x, kxy1 * dx));
    d.y = -(f_fma(dkxy, y, kxy1 * dy));
    d.z = -(f_fma(dkz, z, kz1 * dz) + Q.disk_gm * f_fma(dh, q0, h1 * dq));
    return d;
}

PHX_HDF F3 dforce32(const PotF &Q, float x, float y, float z, float dx, float dy, float dz) {
    float tmax = 0.0f, r2max = 0.0f;
    return dforce32_t<false>(Q, x, y, z, dx, dy, dz, tmax, r2max);
}

// float-float accumulation: (hi, lo) += v
PHX_HDF void ff_add(float &hi, float &lo, float v) {
    const float s = hi + v;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (v - bp);
    const float l = lo + e;
    const float h2 = s + l;
    lo = l - (h2 - s);
    hi = h2;
}

}  // namespace phx
