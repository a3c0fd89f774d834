// FP64 physics of the trace-back path, shared by host and gfx950 device code.
//
// Every function follows the reference operation by operation (cited per function) and the library is
// built with -ffp-contract=off, so host code reproduces the reference's doubles exactly.  Device code
// takes pow, x^1.5, hypot, sin and cos from crmath.h (correctly rounded), so it agrees with the host C
// library wherever that rounds to nearest; it differs only where the C library itself misrounds (glibc:
// about 0.1% of pow arguments, 0.15% of sin / cos, 0.5% of hypot) and in the Box-Muller draws' log / cos,
// which only reach draw covariances and counts.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "crmath.h"
#include "phygas_types.h"

#define PHX_HD __host__ __device__ __forceinline__

namespace phx {

// The transcendentals of the force and the coordinate transform: the C library's on the host (the
// reference's own calls), the correctly rounded ones on the device.  Overloaded on the compilation side,
// so the shared functions below call the right one on each.
__host__ __forceinline__ double pow15(double x) { return ::pow(x, 1.5); }
__device__ __forceinline__ double pow15(double x) { return cr::pow15_cr(x); }
__host__ __forceinline__ void pow_pair(double x, double y1, double y2, double &p1, double &p2) {
    p1 = ::pow(x, y1);
    p2 = ::pow(x, y2);
}
__device__ __forceinline__ void pow_pair(double x, double y1, double y2, double &p1, double &p2) {
    cr::pow2_cr(x, y1, y2, p1, p2);
}
__host__ __forceinline__ double hypot_rn(double x, double y) { return ::hypot(x, y); }
__device__ __forceinline__ double hypot_rn(double x, double y) { return cr::hypot_cr(x, y); }
__host__ __forceinline__ void sin_cos(double x, double &s, double &c) {
    s = ::sin(x);
    c = ::cos(x);
}
__device__ __forceinline__ void sin_cos(double x, double &s, double &c) { cr::sincos_cr(x, s, c); }

constexpr double KMS_PER_KPCMYR = 977.7922216807892;      // reference: src/units.h:14
constexpr double PI = 3.14159265358979323846;
constexpr double STEP_MYR = 0.5;                           // PH_PRE_STEP_MYR
constexpr uint64_t GOLDEN_GAMMA = 0x9E3779B97F4A7C15ULL;   // splitmix64 increment

// Allen & Santillan 1991 parameters in the reference's device order (reference: src/main_phygas_boinc.cpp:79-80;
// values src/potential.c:34-45).
struct Pot {
    double bulge_gm, bulge_b, disk_gm, disk_a, disk_b, halo_gm, halo_a, halo_cut, halo_gamma;
};

inline Pot default_pot() {
    const double G = 4.30091e-6 / (KMS_PER_KPCMYR * KMS_PER_KPCMYR);   // src/units.h:27-28
    Pot p;
    p.bulge_gm = G * (606.0 * 2.325e7);
    p.bulge_b = 0.3873;
    p.disk_gm = G * (3690.0 * 2.325e7);
    p.disk_a = 5.3178;
    p.disk_b = 0.2500;
    p.halo_gm = G * (4615.0 * 2.325e7);
    p.halo_a = 12.0;
    p.halo_cut = 100.0;
    p.halo_gamma = 2.02;
    return p;
}

struct V3 { double x, y, z; };
PHX_HD V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
PHX_HD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
PHX_HD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
PHX_HD V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
PHX_HD double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PHX_HD double norm(V3 a) { return sqrt(dot(a, a)); }

// Acceleration (reference: src/potential.c:54-101).
PHX_HD V3 accel(const Pot &P, V3 p) {
    double R2 = p.x * p.x + p.y * p.y;
    double R = sqrt(R2);
    double r = norm(p);
    V3 a = v3(0.0, 0.0, 0.0);
    double den = pow15(r * r + P.bulge_b * P.bulge_b);
    double fb = -P.bulge_gm / (den > 0 ? den : 1e-12);
    a = a + p * fb;
    double zeta = sqrt(p.z * p.z + P.disk_b * P.disk_b);
    double az_ = P.disk_a + zeta;
    double D2 = R2 + az_ * az_;
    double D = sqrt(D2);
    double D3 = D2 * D;
    double aR = -P.disk_gm * R / D3;
    double azz = -P.disk_gm * p.z * az_ / (zeta > 0 ? zeta : 1e-12) / D3;
    if (R > 1e-12) {
        a.x += aR * p.x / R;
        a.y += aR * p.y / R;
    }
    a.z += azz;
    double rs = r > 1e-12 ? r : 1e-12;
    double re = rs < P.halo_cut ? rs : P.halo_cut;
    double xh = re / P.halo_a;
    double pg, pg1;   // xh^gamma, xh^(gamma - 1)
    pow_pair(xh, P.halo_gamma, P.halo_gamma - 1.0, pg, pg1);
    double gm = P.halo_gm * pg / (1.0 + pg1);
    double fh = -gm / (rs * rs * rs);
    return a + p * fh;
}

// Kick-drift-kick with the acceleration carried over (reference: src/phygas.c:85-96).
struct Body { V3 pos, vel, acc; };
PHX_HD Body body_start(const Pot &P, V3 pos, V3 vel) { Body b; b.pos = pos; b.vel = vel; b.acc = accel(P, pos); return b; }
PHX_HD void kdk(const Pot &P, Body &b, double dt) {
    b.vel = b.vel + b.acc * (0.5 * dt);
    b.pos = b.pos + b.vel * dt;
    b.acc = accel(P, b.pos);
    b.vel = b.vel + b.acc * (0.5 * dt);
}

// splitmix64 and Box-Muller (reference: src/montecarlo.h:26-41, src/montecarlo.c:4-17).
PHX_HD uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
PHX_HD uint64_t rng_next(uint64_t &st) { st += GOLDEN_GAMMA; return mix64(st); }
PHX_HD double to_unit(uint64_t u) { return (double)(u >> 11) * (1.0 / 9007199254740992.0); }
PHX_HD double box_muller(double u1, double u2) {
    if (u1 <= 0.0) u1 = 1.0 / 18446744073709551616.0;
    double r = sqrt(-2.0 * log(u1));
    return r * cos(2.0 * PI * u2);
}
PHX_HD double gaussian(uint64_t &st) {
    double u1 = to_unit(rng_next(st));
    double u2 = to_unit(rng_next(st));
    return box_muller(u1, u2);
}

// ICRS -> Galactic rotation and the Sun (reference: src/coords.c:31-35, 58-65).
PHX_HD V3 rot_g(V3 v) {
    return v3(-0.0548755604162154 * v.x + -0.8734370902348850 * v.y + -0.4838350155487132 * v.z,
              0.4941094278755837 * v.x + -0.4448296299600112 * v.y + 0.7469822444972189 * v.z,
              -0.8676661490190047 * v.x + -0.1980763734312015 * v.y + 0.4559837761750669 * v.z);
}
PHX_HD V3 gc_pos(V3 h) { V3 g = rot_g(h); return v3(g.x - 8.122, g.y, g.z + 0.0208); }
PHX_HD V3 gc_vel(V3 h) {
    V3 g = rot_g(h);
    return g + v3(11.1 / KMS_PER_KPCMYR, (229.0 + 12.24) / KMS_PER_KPCMYR, 7.25 / KMS_PER_KPCMYR);
}

struct Astro { double ra, dec, plx, pmra, pmdec, rv; };

// Astrometry -> Galactocentric phase space (reference: src/coords.c:77-104).
PHX_HD void astro_to_gc(const Astro &a, V3 &pos, V3 &vel) {
    double ra = a.ra * (PI / 180.0), dec = a.dec * (PI / 180.0);
    double dk = 1.0 / a.plx, dp = dk * 1000.0;
    double sa, ca, sd, cd;
    sin_cos(ra, sa, ca);
    sin_cos(dec, sd, cd);
    V3 rh = v3(cd * ca, cd * sa, sd);
    V3 eh = v3(-sa, ca, 0.0);
    V3 nh = v3(-sd * ca, -sd * sa, cd);
    pos = gc_pos(rh * dk);
    const double k = (1.495978707e8 / 3.15576e7) * 1e-3;
    double vra = k * a.pmra * dp, vde = k * a.pmdec * dp;
    V3 vk = (rh * a.rv + eh * vra) + nh * vde;
    vel = gc_vel(v3(vk.x / KMS_PER_KPCMYR, vk.y / KMS_PER_KPCMYR, vk.z / KMS_PER_KPCMYR));
}

// Radial velocity of circular Galactic rotation at pos, seen from the Sun (reference: src/phygas.c:25-36).
PHX_HD double rv_rotation(const Pot &P, V3 p) {
    double R = hypot_rn(p.x, p.y);
    if (R <= 0) return 0.0;
    V3 a = accel(P, v3(p.x, p.y, 0.0));
    double ar = -(a.x * p.x + a.y * p.y) / R;
    double vc = ar > 0 ? sqrt(R * ar) : 0.0;
    V3 v = v3(vc * p.y / R, -vc * p.x / R, 0.0);
    V3 sun_p = gc_pos(v3(0.0, 0.0, 0.0)), sun_v = gc_vel(v3(0.0, 0.0, 0.0));
    V3 los = p - sun_p;
    double n = norm(los);
    if (n <= 0) return 0.0;
    return dot(v - sun_v, los * (1.0 / n)) * KMS_PER_KPCMYR;
}

PHX_HD double draw(uint64_t *st) { return st ? gaussian(*st) : 0.0; }

// RV from the measurement or the rotation prior (reference: src/phygas.c:38-48).
PHX_HD void with_rv(const Pot &P, Astro a, int has_rv, double rv, double e_rv, double prior, uint64_t *st, V3 &pos,
                    V3 &vel) {
    if (has_rv) {
        a.rv = rv + e_rv * draw(st);
    } else {
        a.rv = 0.0;
        V3 p0, v0;
        astro_to_gc(a, p0, v0);
        a.rv = rv_rotation(P, p0) + prior * draw(st);
    }
    astro_to_gc(a, pos, vel);
}

// One draw (st == nullptr: the nominal); 0 or -1 (reference: src/phygas.c:50-66).
PHX_HD int sample_target(const Pot &P, const PhTarget &t, uint64_t *st, V3 &pos, V3 &vel) {
    Astro a;
    a.ra = t.ra; a.dec = t.dec;
    for (int tries = 0; tries < 50; tries++) {
        double plx = t.plx > 0 ? t.plx + t.e_plx * draw(st) : 1000.0 / (t.dist_pc * exp(t.e_lndist * draw(st)));
        if (!(plx > 0)) continue;
        a.plx = plx;
        a.pmra = t.pmra + t.e_pmra * draw(st);
        a.pmdec = t.pmdec + t.e_pmdec * draw(st);
        with_rv(P, a, t.has_rv, t.rv, t.e_rv, t.kind == 'P' ? PH_PULSAR_RV_PRIOR : PH_STAR_RV_PRIOR, st, pos, vel);
        return 0;
    }
    return -1;
}

// (reference: src/phygas.c:68-82)
PHX_HD int sample_cluster(const Pot &P, const PhCluster &c, uint64_t *st, V3 &pos, V3 &vel) {
    Astro a;
    a.ra = c.ra; a.dec = c.dec;
    for (int tries = 0; tries < 50; tries++) {
        double d = c.dist_pc + c.e_dist_pc * draw(st);
        if (!(d > 0)) continue;
        a.plx = 1000.0 / d;
        a.pmra = c.pmra + c.e_pmra * draw(st);
        a.pmdec = c.pmdec + c.e_pmdec * draw(st);
        with_rv(P, a, c.has_rv, c.rv, c.e_rv, PH_CLUSTER_RV_PRIOR, st, pos, vel);
        return 0;
    }
    return -1;
}

// m^T M^-1 m of a symmetric 3x3 (xx xy xz yy yz zz); -1 if not positive (reference: src/phygas.c:348-356).
PHX_HD double mahalanobis(const double M[6], double mx, double my, double mz) {
    double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5];
    double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    double det = a * c00 + b * c01 + c * c02;
    if (!(det > 0)) return -1;
    double c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
    return (mx * (c00 * mx + c01 * my + c02 * mz) + my * (c01 * mx + c11 * my + c12 * mz) +
            mz * (c02 * mx + c12 * my + c22 * mz)) / det;
}

// Where on the segment r0 -> r1 the distance from the origin is smallest, s in [0,1]
// (reference: src/phygas.c:147-156).
PHX_HD double seg_param(V3 r0, V3 r1) {
    V3 dr = r1 - r0;
    double dd = dot(dr, dr), s = 0;
    if (dd > 0) {
        s = -dot(r0, dr) / dd;
        if (s < 0) s = 0; else if (s > 1) s = 1;
    }
    return s;
}

// 64-bit FNV-style hash with the reference's basis (reference: src/phygas.c:452-458), in two halves so the
// device can finish a per-target prefix with a cluster name.
constexpr uint64_t FNV_BASIS = 1469598103934665603ULL, FNV_PRIME = 1099511628211ULL;
PHX_HD uint64_t fnv_feed(uint64_t h, const char *s, int maxlen) {
    for (int i = 0; i < maxlen && s[i]; i++) h = (h ^ (unsigned char)s[i]) * FNV_PRIME;
    return h;
}
PHX_HD uint64_t seed_prefix(const char *a) { return (fnv_feed(FNV_BASIS, a, 1 << 30) ^ (unsigned char)'|') * FNV_PRIME; }

// 512 Cholesky draws around m: counts inside r50 and rtot (reference: src/phygas.c:427-443).
PHX_HD void finish_counts(const double C[6], V3 m, uint64_t seed, double r50, double rt, int &in50, int &intot) {
    double l00 = sqrt(fmax(C[0], 0.0)), l10 = l00 > 0 ? C[1] / l00 : 0, l20 = l00 > 0 ? C[2] / l00 : 0;
    double l11 = sqrt(fmax(C[3] - l10 * l10, 0.0)), l21 = l11 > 0 ? (C[4] - l20 * l10) / l11 : 0;
    double l22 = sqrt(fmax(C[5] - l20 * l20 - l21 * l21, 0.0));
    uint64_t st = seed;
    int a = 0, b = 0;
    for (int i = 0; i < 512; i++) {
        double z0 = gaussian(st), z1 = gaussian(st), z2 = gaussian(st);
        double x = m.x + l00 * z0, y = m.y + l10 * z0 + l11 * z1, z = m.z + l20 * z0 + l21 * z1 + l22 * z2;
        double r = sqrt(x * x + y * y + z * z);
        if (r < r50) a++;
        if (r < rt) b++;
    }
    in50 = a;
    intot = b;
}

}  // namespace phx
