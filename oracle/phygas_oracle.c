/* ORACLE -- test infrastructure only (the checker; never the thing measured or shipped).
 *
 * CPU FP64 restatement of the Phygas trace-back hot path, written from the reference's semantics
 * operation by operation so that, built with gcc -O2 -std=c99 (no FMA contraction, no fast-math) on
 * glibc, every number is bit-identical to the compiled reference.  Each function cites the reference
 * lines it follows.  Pinned by tests/test_oracle.py against the golden vectors in tests/golden/ (made
 * by the compiled reference) and against closed forms / brute force where the reference leaves gaps.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library. */
#define _POSIX_C_SOURCE 200809L
#include "phygas_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef ORC_ULP_SHIM
/* Test-only build (liboracle_ulp.so): pow, sin, cos and hypot of the C library moved by one unit in the last
 * place in a chosen share of calls (orc_ulp_set; a hash of the arguments decides, so a given argument always
 * gets the same result, as under another libm).  It measures how far last-bit differences in the
 * transcendentals move the tracks and the chi2 minima. */
double orc_ulp_pow(double x, double y);
double orc_ulp_sin(double x);
double orc_ulp_cos(double x);
double orc_ulp_hypot(double x, double y);
#define pow orc_ulp_pow
#define sin orc_ulp_sin
#define cos orc_ulp_cos
#define hypot orc_ulp_hypot
#endif

/* ---- units and constants (reference: src/units.h:14,27-28; src/coords.c:18,31-35,58-65,71) ---- */
#define KMS_PER_KPCMYR 977.7922216807892
#define G_INTERNAL (4.30091e-6 / (KMS_PER_KPCMYR * KMS_PER_KPCMYR))
#define MGAL 2.325e7
#define PI_D 3.14159265358979323846
#define DEG2RAD (PI_D / 180.0)
#define RAD2DEG (180.0 / PI_D)
#define AUYR_KMS (1.495978707e8 / 3.15576e7)

static const double ROT_G[3][3] = {
    {-0.0548755604162154, -0.8734370902348850, -0.4838350155487132},
    {0.4941094278755837, -0.4448296299600112, 0.7469822444972189},
    {-0.8676661490190047, -0.1980763734312015, 0.4559837761750669},
};
static const double SUN_R0 = 8.122, SUN_Z0 = 0.0208, SUN_VC = 229.0, SUN_U = 11.1, SUN_V = 12.24, SUN_W = 7.25;

/* Allen & Santillan 1991 parameters (reference: src/potential.c:34-45); device order of
 * src/main_phygas_boinc.cpp:79-80. */
static const double MW_BULGE_GM = G_INTERNAL * (606.0 * MGAL), MW_BULGE_B = 0.3873;
static const double MW_DISK_GM = G_INTERNAL * (3690.0 * MGAL), MW_DISK_A = 5.3178, MW_DISK_B = 0.2500;
static const double MW_HALO_GM = G_INTERNAL * (4615.0 * MGAL), MW_HALO_A = 12.0, MW_HALO_CUT = 100.0,
                    MW_HALO_GAMMA = 2.02;

void orc_params(double p[9]) {
    p[0] = MW_BULGE_GM; p[1] = MW_BULGE_B; p[2] = MW_DISK_GM; p[3] = MW_DISK_A; p[4] = MW_DISK_B;
    p[5] = MW_HALO_GM; p[6] = MW_HALO_A; p[7] = MW_HALO_CUT; p[8] = MW_HALO_GAMMA;
}

/* ---- vectors (reference: src/vec3.h) ---- */
static inline Vec3 mk(double x, double y, double z) { Vec3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline Vec3 vadd(Vec3 a, Vec3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline Vec3 vsub(Vec3 a, Vec3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline Vec3 vmul(Vec3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double vdot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double vlen(Vec3 a) { return sqrt(vdot(a, a)); }

/* ---- potential (reference: src/potential.c:54-101 acceleration, :103-136 energy) ---- */
static double halo_gm(double r) {   /* potential.c:54-58 */
    double re = r < MW_HALO_CUT ? r : MW_HALO_CUT;
    double x = re / MW_HALO_A;
    return MW_HALO_GM * pow(x, MW_HALO_GAMMA) / (1.0 + pow(x, MW_HALO_GAMMA - 1.0));
}

static Vec3 accel(Vec3 p) {   /* potential.c:60-101 */
    double R2 = p.x * p.x + p.y * p.y;
    double R = sqrt(R2);
    double r = vlen(p);
    Vec3 a = mk(0, 0, 0);
    double den = pow(r * r + MW_BULGE_B * MW_BULGE_B, 1.5);
    double fb = -MW_BULGE_GM / (den > 0 ? den : 1e-12);
    a = vadd(a, vmul(p, fb));
    double zeta = sqrt(p.z * p.z + MW_DISK_B * MW_DISK_B);
    double az_ = MW_DISK_A + zeta;
    double D2 = R2 + az_ * az_;
    double D = sqrt(D2);
    double D3 = D2 * D;
    double aR = -MW_DISK_GM * R / D3;
    double az = -MW_DISK_GM * p.z * az_ / (zeta > 0 ? zeta : 1e-12) / D3;
    if (R > 1e-12) {
        a.x += aR * p.x / R;
        a.y += aR * p.y / R;
    }
    a.z += az;
    double rs = r > 1e-12 ? r : 1e-12;
    double fh = -halo_gm(rs) / (rs * rs * rs);
    a = vadd(a, vmul(p, fh));
    return a;
}

static void pot_accel(const Potential *pot, Vec3 pos, Vec3 *out) { (void)pot; *out = accel(pos); }
static double pot_energy(const Potential *pot, Vec3 pos) {
    (void)pot;
    return orc_energy(pos.x, pos.y, pos.z);
}
static const Potential g_potential = {pot_accel, pot_energy, NULL};
const Potential *ph_potential(void) { return &g_potential; }

void orc_accel(double x, double y, double z, double out[3]) {
    Vec3 a = accel(mk(x, y, z));
    out[0] = a.x; out[1] = a.y; out[2] = a.z;
}

double orc_energy(double x, double y, double z) {   /* potential.c:103-136 */
    double R2 = x * x + y * y, r = sqrt(x * x + y * y + z * z);
    double pb = -MW_BULGE_GM / sqrt(r * r + MW_BULGE_B * MW_BULGE_B);
    double zeta = sqrt(z * z + MW_DISK_B * MW_DISK_B), az_ = MW_DISK_A + zeta;
    double pd = -MW_DISK_GM / sqrt(R2 + az_ * az_);
    double rs = r > 1e-12 ? r : 1e-12, gmL = halo_gm(MW_HALO_CUT), ph;
    if (rs >= MW_HALO_CUT) {
        ph = -gmL / rs;
    } else {
        double g1 = MW_HALO_GAMMA - 1.0, xL = MW_HALO_CUT / MW_HALO_A, xr = rs / MW_HALO_A;
        double integ = (1.0 / g1) * (log(1.0 + pow(xL, g1)) - log(1.0 + pow(xr, g1)));
        ph = -(MW_HALO_GM / MW_HALO_A) * integ - gmL / MW_HALO_CUT;
    }
    return pb + pd + ph;
}

/* ---- RNG (reference: src/montecarlo.h:26-41, src/montecarlo.c:4-17) ---- */
static inline uint64_t rng_u64(McRng *g) {
    uint64_t z = (g->state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double rng_unit(McRng *g) { return (double)(rng_u64(g) >> 11) * (1.0 / 9007199254740992.0); }

void mc_rng_seed(McRng *g, uint64_t seed) { g->state = seed; }
uint64_t mc_rng_next_u64(McRng *g) { return rng_u64(g); }
double mc_rng_uniform(McRng *g) { return rng_unit(g); }

double mc_gaussian(McRng *g) {
    double u1 = rng_unit(g);
    if (u1 <= 0.0) u1 = 1.0 / 18446744073709551616.0;
    double u2 = rng_unit(g);
    double r = sqrt(-2.0 * log(u1));
    return r * cos(2.0 * PI_D * u2);
}
static double gauss_or_zero(McRng *g) { return g ? mc_gaussian(g) : 0.0; }   /* phygas.c:21 */

/* ---- coordinates (reference: src/coords.c:37-51, 73-161) ---- */
static Vec3 rot(const double m[3][3], Vec3 v) {
    return mk(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
              m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}
static Vec3 rot_t(const double m[3][3], Vec3 v) {
    return mk(m[0][0] * v.x + m[1][0] * v.y + m[2][0] * v.z, m[0][1] * v.x + m[1][1] * v.y + m[2][1] * v.z,
              m[0][2] * v.x + m[1][2] * v.y + m[2][2] * v.z);
}
static Vec3 helio_pos(Vec3 p) {   /* coords.c:106-113 */
    Vec3 h = rot(ROT_G, p);
    return mk(h.x - SUN_R0, h.y, h.z + SUN_Z0);
}
static Vec3 helio_vel(Vec3 v) {   /* coords.c:115-124 */
    Vec3 h = rot(ROT_G, v);
    Vec3 s = mk(SUN_U / KMS_PER_KPCMYR, (SUN_VC + SUN_V) / KMS_PER_KPCMYR, SUN_W / KMS_PER_KPCMYR);
    return vadd(h, s);
}

typedef struct { double ra, dec, plx, pmra, pmdec, rv; } Astro;

static void astro_to_gc(Astro a, PhaseSpace *out) {   /* coords.c:77-104 */
    double ra = a.ra * DEG2RAD, dec = a.dec * DEG2RAD;
    double dk = 1.0 / a.plx, dp = dk * 1000.0;
    Vec3 rh = mk(cos(dec) * cos(ra), cos(dec) * sin(ra), sin(dec));
    Vec3 eh = mk(-sin(ra), cos(ra), 0.0);
    Vec3 nh = mk(-sin(dec) * cos(ra), -sin(dec) * sin(ra), cos(dec));
    out->pos = helio_pos(vmul(rh, dk));
    double vra = AUYR_KMS * 1e-3 * a.pmra * dp;
    double vde = AUYR_KMS * 1e-3 * a.pmdec * dp;
    Vec3 vk = vadd(vadd(vmul(rh, a.rv), vmul(eh, vra)), vmul(nh, vde));
    out->vel = helio_vel(mk(vk.x / KMS_PER_KPCMYR, vk.y / KMS_PER_KPCMYR, vk.z / KMS_PER_KPCMYR));
}

void orc_astrometry_to_gc(double ra, double dec, double plx, double pmra, double pmdec, double rv, double out[6]) {
    Astro a = {ra, dec, plx, pmra, pmdec, rv};
    PhaseSpace p;
    astro_to_gc(a, &p);
    out[0] = p.pos.x; out[1] = p.pos.y; out[2] = p.pos.z; out[3] = p.vel.x; out[4] = p.vel.y; out[5] = p.vel.z;
}

void orc_gc_to_astrometry(const double ps[6], double out[6]) {   /* coords.c:126-161 */
    Vec3 hg = mk(ps[0] + SUN_R0, ps[1], ps[2] - SUN_Z0);
    double dk = vlen(hg), dp = dk * 1000.0;
    Vec3 rh = rot_t(ROT_G, vmul(hg, 1.0 / dk));
    double ra = atan2(rh.y, rh.x), dec = asin(rh.z);
    if (ra < 0) ra += 2.0 * PI_D;
    Vec3 eh = mk(-sin(ra), cos(ra), 0.0), nh = mk(-sin(dec) * cos(ra), -sin(dec) * sin(ra), cos(dec));
    Vec3 vs = mk(SUN_U, SUN_VC + SUN_V, SUN_W);
    Vec3 vg = mk(ps[3] * KMS_PER_KPCMYR, ps[4] * KMS_PER_KPCMYR, ps[5] * KMS_PER_KPCMYR);
    Vec3 vi = rot_t(ROT_G, vsub(vg, vs));
    out[0] = ra * RAD2DEG; out[1] = dec * RAD2DEG; out[2] = 1.0 / dk;
    out[3] = vdot(vi, eh) / (AUYR_KMS * 1e-3 * dp);
    out[4] = vdot(vi, nh) / (AUYR_KMS * 1e-3 * dp);
    out[5] = vdot(vi, rh);
}

/* ---- draws (reference: src/phygas.c:12-82) ---- */
static Vec3 g_sun_pos, g_sun_vel;
static int g_ready;

void ph_init(void) {   /* phygas.c:12-17 */
    g_sun_pos = helio_pos(mk(0, 0, 0));
    g_sun_vel = helio_vel(mk(0, 0, 0));
    g_ready = 1;
}
static void need_init(void) { if (!g_ready) ph_init(); }

static double rv_rotation(Vec3 p) {   /* phygas.c:25-36 */
    double R = hypot(p.x, p.y);
    if (R <= 0) return 0.0;
    Vec3 a = accel(mk(p.x, p.y, 0.0));
    double ar = -(a.x * p.x + a.y * p.y) / R;
    double vc = ar > 0 ? sqrt(R * ar) : 0.0;
    Vec3 v = mk(vc * p.y / R, -vc * p.x / R, 0.0);
    Vec3 los = vsub(p, g_sun_pos);
    double n = vlen(los);
    if (n <= 0) return 0.0;
    return vdot(vsub(v, g_sun_vel), vmul(los, 1.0 / n)) * KMS_PER_KPCMYR;
}

static void with_rv(Astro a, int has_rv, double rv, double e_rv, double prior, McRng *g, PhaseSpace *out) {
    /* phygas.c:38-48 */
    if (has_rv) {
        a.rv = rv + e_rv * gauss_or_zero(g);
    } else {
        a.rv = 0.0;
        PhaseSpace p;
        astro_to_gc(a, &p);
        a.rv = rv_rotation(p.pos) + prior * gauss_or_zero(g);
    }
    astro_to_gc(a, out);
}

int ph_sample_target(const PhTarget *t, McRng *g, PhaseSpace *out) {   /* phygas.c:50-66 */
    need_init();
    Astro a;
    a.ra = t->ra; a.dec = t->dec;
    for (int tries = 0; tries < 50; tries++) {
        double plx = t->plx > 0 ? t->plx + t->e_plx * gauss_or_zero(g)
                                : 1000.0 / (t->dist_pc * exp(t->e_lndist * gauss_or_zero(g)));
        if (!(plx > 0)) continue;
        a.plx = plx;
        a.pmra = t->pmra + t->e_pmra * gauss_or_zero(g);
        a.pmdec = t->pmdec + t->e_pmdec * gauss_or_zero(g);
        with_rv(a, t->has_rv, t->rv, t->e_rv, t->kind == 'P' ? PH_PULSAR_RV_PRIOR : PH_STAR_RV_PRIOR, g, out);
        return 0;
    }
    return -1;
}

int ph_sample_cluster(const PhCluster *c, McRng *g, PhaseSpace *out) {   /* phygas.c:68-82 */
    need_init();
    Astro a;
    a.ra = c->ra; a.dec = c->dec;
    for (int tries = 0; tries < 50; tries++) {
        double d = c->dist_pc + c->e_dist_pc * gauss_or_zero(g);
        if (!(d > 0)) continue;
        a.plx = 1000.0 / d;
        a.pmra = c->pmra + c->e_pmra * gauss_or_zero(g);
        a.pmdec = c->pmdec + c->e_pmdec * gauss_or_zero(g);
        with_rv(a, c->has_rv, c->rv, c->e_rv, PH_CLUSTER_RV_PRIOR, g, out);
        return 0;
    }
    return -1;
}

/* ---- leapfrog (reference: src/phygas.c:85-96) ---- */
typedef struct { Vec3 pos, vel, acc; } Orbit;
static void orbit_start(Orbit *b, PhaseSpace ps) { b->pos = ps.pos; b->vel = ps.vel; b->acc = accel(ps.pos); }
static inline void kdk(Orbit *b, double dt) {
    b->vel = vadd(b->vel, vmul(b->acc, 0.5 * dt));
    b->pos = vadd(b->pos, vmul(b->vel, dt));
    b->acc = accel(b->pos);
    b->vel = vadd(b->vel, vmul(b->acc, 0.5 * dt));
}

void orc_body_step(double b[9], double dt) {
    Orbit o = {mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), mk(b[6], b[7], b[8])};
    kdk(&o, dt);
    b[0] = o.pos.x; b[1] = o.pos.y; b[2] = o.pos.z; b[3] = o.vel.x; b[4] = o.vel.y; b[5] = o.vel.z;
    b[6] = o.acc.x; b[7] = o.acc.y; b[8] = o.acc.z;
}

/* ---- prefilter tracks (reference: src/phygas.c:98-168) ---- */
static void nominal_path(PhaseSpace ps, double tmax, PhTrack *out) {   /* phygas.c:98-110 */
    int n = (int)ceil(tmax / PH_PRE_STEP_MYR) + 1;
    out->n = n;
    out->pos = malloc(sizeof(Vec3) * n);
    Orbit b;
    orbit_start(&b, ps);
    double dt = -PH_PRE_STEP_MYR / 10;
    out->pos[0] = b.pos;
    for (int k = 1; k < n; k++) {
        for (int i = 0; i < 10; i++) kdk(&b, dt);
        out->pos[k] = b.pos;
    }
}

static void path_spread(PhaseSpace nom, const PhaseSpace *d, int n, PhTrack *out) {   /* phygas.c:112-120 */
    double sp = 0, sv = 0;
    for (int i = 0; i < n; i++) {
        Vec3 dp = vsub(d[i].pos, nom.pos), dv = vsub(d[i].vel, nom.vel);
        sp += vdot(dp, dp); sv += vdot(dv, dv);
    }
    out->sig_pos = n ? sqrt(sp / n) : 0;
    out->sig_vel = n ? sqrt(sv / n) : 0;
}

void ph_track_target(const PhTarget *t, double tmax, int n_spread, uint64_t seed, PhTrack *out) {
    PhaseSpace nom, d[64];
    ph_sample_target(t, NULL, &nom);
    McRng g = {seed};
    if (n_spread > 64) n_spread = 64;
    int k = 0;
    for (int i = 0; i < n_spread; i++) if (ph_sample_target(t, &g, &d[k]) == 0) k++;
    nominal_path(nom, tmax, out);
    path_spread(nom, d, k, out);
}

void ph_track_cluster(const PhCluster *c, int n_spread, uint64_t seed, PhTrack *out) {
    PhaseSpace nom, d[64];
    ph_sample_cluster(c, NULL, &nom);
    McRng g = {seed};
    if (n_spread > 64) n_spread = 64;
    int k = 0;
    for (int i = 0; i < n_spread; i++) if (ph_sample_cluster(c, &g, &d[k]) == 0) k++;
    nominal_path(nom, c->tmax_myr, out);
    path_spread(nom, d, k, out);
}

void ph_track_free(PhTrack *t) { free(t->pos); t->pos = NULL; t->n = 0; }

static double segment_closest(Vec3 r0, Vec3 r1, double *s_out) {   /* phygas.c:147-156 */
    Vec3 dr = vsub(r1, r0);
    double dd = vdot(dr, dr), s = 0;
    if (dd > 0) {
        s = -vdot(r0, dr) / dd;
        if (s < 0) s = 0; else if (s > 1) s = 1;
    }
    *s_out = s;
    return vlen(vadd(r0, vmul(dr, s)));
}

double orc_seg_min(const double r0[3], const double r1[3], double *s) {
    return segment_closest(mk(r0[0], r0[1], r0[2]), mk(r1[0], r1[1], r1[2]), s);
}

int ph_prefilter(const PhTrack *tt, const PhTrack *ct, const PhCluster *c) {   /* phygas.c:158-168 */
    int n = tt->n < ct->n ? tt->n : ct->n;
    double sp = hypot(tt->sig_pos, ct->sig_pos), sv = hypot(tt->sig_vel, ct->sig_vel);
    double r0 = (3.0 * c->rtot_pc + 10.0) / 1000.0;
    for (int k = 0; k + 1 < n; k++) {
        double s, d = segment_closest(vsub(tt->pos[k], ct->pos[k]), vsub(tt->pos[k + 1], ct->pos[k + 1]), &s);
        double t = (k + s) * PH_PRE_STEP_MYR;
        if (d < r0 + 3.0 * (sp + sv * t)) return 1;
    }
    return 0;
}

/* ---- Monte Carlo pairs (reference: src/phygas.c:170-249) ---- */
static int by_value(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

static double interp_quantile(const double *v, int n, double q) {   /* phygas.c:175-180 */
    double pos = q * (n - 1);
    int i = (int)pos;
    if (i >= n - 1) return v[n - 1];
    return v[i] + (pos - i) * (v[i + 1] - v[i]);
}

int ph_pair_draw(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, double *ps,
                 int *steps_out, double *dt_out) {   /* phygas.c:182-198 */
    McRng g = {seed};
    int steps = (int)ceil(c->tmax_myr / dt_myr);
    *steps_out = steps;
    *dt_out = -c->tmax_myr / steps;
    int m = 0;
    for (int k = 0; k < n_clones; k++) {
        PhaseSpace a, b;
        if (ph_sample_target(t, &g, &a) || ph_sample_cluster(c, &g, &b)) continue;
        double *o = ps + (size_t)m * 12;
        o[0] = a.pos.x; o[1] = a.pos.y; o[2] = a.pos.z; o[3] = a.vel.x; o[4] = a.vel.y; o[5] = a.vel.z;
        o[6] = b.pos.x; o[7] = b.pos.y; o[8] = b.pos.z; o[9] = b.vel.x; o[10] = b.vel.y; o[11] = b.vel.z;
        m++;
    }
    return m;
}

void ph_pair_clone(const double *ps, int steps, double dt, double *best_out, double *tbest_out,
                   double *vbest_out) {   /* phygas.c:200-216 */
    PhaseSpace pt = {{ps[0], ps[1], ps[2]}, {ps[3], ps[4], ps[5]}}, pc = {{ps[6], ps[7], ps[8]}, {ps[9], ps[10], ps[11]}};
    Orbit a, b;
    orbit_start(&a, pt);
    orbit_start(&b, pc);
    Vec3 r0 = vsub(a.pos, b.pos);
    double best = vlen(r0), tbest = 0, vbest = vlen(vsub(a.vel, b.vel));
    for (int i = 1; i <= steps; i++) {
        kdk(&a, dt);
        kdk(&b, dt);
        Vec3 r1 = vsub(a.pos, b.pos);
        double s, dm = segment_closest(r0, r1, &s);
        if (dm < best) {
            best = dm;
            tbest = (i - 1 + s) * -dt;
            vbest = vlen(vsub(a.vel, b.vel));
        }
        r0 = r1;
    }
    *best_out = best; *tbest_out = tbest; *vbest_out = vbest;
}

void ph_pair_reduce(const PhCluster *c, int m, const double *best, const double *tbest, const double *vbest,
                    PhResult *res) {   /* phygas.c:218-238 */
    int cap = m > 0 ? m : 1;
    double *d = malloc(sizeof(double) * cap * 3), *tt = d + cap, *vv = d + 2 * cap;
    int in50 = 0, intot = 0;
    for (int k = 0; k < m; k++) {
        d[k] = best[k] * 1000.0;
        tt[k] = tbest[k];
        vv[k] = vbest[k] * KMS_PER_KPCMYR;
        if (d[k] < c->r50_pc) in50++;
        if (d[k] < c->rtot_pc) intot++;
    }
    memset(res, 0, sizeof(*res));
    res->n = m;
    if (m > 0) {
        res->f_r50 = (double)in50 / m;
        res->f_rtot = (double)intot / m;
        qsort(d, m, sizeof(double), by_value);
        qsort(tt, m, sizeof(double), by_value);
        qsort(vv, m, sizeof(double), by_value);
        res->d05 = interp_quantile(d, m, 0.05); res->d50 = interp_quantile(d, m, 0.5);
        res->d95 = interp_quantile(d, m, 0.95);
        res->t05 = interp_quantile(tt, m, 0.05); res->t50 = interp_quantile(tt, m, 0.5);
        res->t95 = interp_quantile(tt, m, 0.95);
        res->v05 = interp_quantile(vv, m, 0.05); res->v50 = interp_quantile(vv, m, 0.5);
        res->v95 = interp_quantile(vv, m, 0.95);
    }
    free(d);
}

void ph_pair(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, PhResult *res) {
    int cap = n_clones > 0 ? n_clones : 1;   /* phygas.c:240-249 */
    double *ps = malloc(sizeof(double) * 12 * cap), *out = malloc(sizeof(double) * 3 * cap);
    int steps;
    double dt;
    int m = ph_pair_draw(t, c, n_clones, dt_myr, seed, ps, &steps, &dt);
    for (int k = 0; k < m; k++) ph_pair_clone(ps + (size_t)k * 12, steps, dt, &out[k], &out[m + k], &out[2 * m + k]);
    ph_pair_reduce(c, m, out, out + m, out + 2 * m, res);
    free(ps);
    free(out);
}

/* ---- linear tracks (reference: src/phygas.c:252-345) ---- */
static void draw_cov(PhaseSpace nom, const PhaseSpace *d, int nd, double sig[6][6]) {   /* phygas.c:256-264 */
    memset(sig, 0, sizeof(double[6][6]));
    for (int i = 0; i < nd; i++) {
        double e[6] = {d[i].pos.x - nom.pos.x, d[i].pos.y - nom.pos.y, d[i].pos.z - nom.pos.z,
                       d[i].vel.x - nom.vel.x, d[i].vel.y - nom.vel.y, d[i].vel.z - nom.vel.z};
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) sig[a][b] += e[a] * e[b];
    }
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) sig[a][b] /= nd > 0 ? nd : 1;
}

static void stm_track(PhaseSpace nom, double sig[6][6], double tmax, PhLin *out) {   /* phygas.c:286-323 */
    Orbit o[7];
    orbit_start(&o[0], nom);
    for (int j = 0; j < 6; j++) {
        PhaseSpace p = nom;
        double e = j < 3 ? PH_LIN_EPS_POS : PH_LIN_EPS_VEL;
        switch (j) {
            case 0: p.pos.x += e; break;
            case 1: p.pos.y += e; break;
            case 2: p.pos.z += e; break;
            case 3: p.vel.x += e; break;
            case 4: p.vel.y += e; break;
            default: p.vel.z += e; break;
        }
        orbit_start(&o[j + 1], p);
    }
    int n = (int)ceil(tmax / PH_PRE_STEP_MYR) + 1;
    out->n = n;
    out->pos = malloc(sizeof(Vec3) * n);
    out->vel = malloc(sizeof(Vec3) * n);
    out->S = malloc(sizeof(double[6]) * n);
    double dt = -PH_PRE_STEP_MYR / PH_LIN_SUB;
    for (int k = 0; k < n; k++) {
        if (k > 0)
            for (int i = 0; i < PH_LIN_SUB; i++)
                for (int j = 0; j < 7; j++) kdk(&o[j], dt);
        out->pos[k] = o[0].pos;
        out->vel[k] = o[0].vel;
        double A[3][6];
        for (int j = 0; j < 6; j++) {
            double e = j < 3 ? PH_LIN_EPS_POS : PH_LIN_EPS_VEL;
            Vec3 d = vsub(o[j + 1].pos, o[0].pos);
            A[0][j] = d.x / e; A[1][j] = d.y / e; A[2][j] = d.z / e;
        }
        double AS[3][6] = {{0}};
        for (int r = 0; r < 3; r++)
            for (int b = 0; b < 6; b++)
                for (int a = 0; a < 6; a++) AS[r][b] += A[r][a] * sig[a][b];
        double S[3][3] = {{0}};
        for (int r = 0; r < 3; r++)
            for (int s = 0; s < 3; s++)
                for (int b = 0; b < 6; b++) S[r][s] += AS[r][b] * A[s][b];
        out->S[k][0] = S[0][0]; out->S[k][1] = S[0][1]; out->S[k][2] = S[0][2];
        out->S[k][3] = S[1][1]; out->S[k][4] = S[1][2]; out->S[k][5] = S[2][2];
    }
}

void ph_lin_target_init(const PhTarget *t, int n_samp, uint64_t seed, double ps0[6], double sig[6][6]) {
    PhaseSpace nom, *d = malloc(sizeof(PhaseSpace) * (n_samp > 0 ? n_samp : 1));   /* phygas.c:274-284 */
    ph_sample_target(t, NULL, &nom);
    McRng g = {seed};
    int k = 0;
    for (int i = 0; i < n_samp; i++) if (ph_sample_target(t, &g, &d[k]) == 0) k++;
    draw_cov(nom, d, k, sig);
    free(d);
    ps0[0] = nom.pos.x; ps0[1] = nom.pos.y; ps0[2] = nom.pos.z;
    ps0[3] = nom.vel.x; ps0[4] = nom.vel.y; ps0[5] = nom.vel.z;
}

void ph_lin_target(const PhTarget *t, double tmax, int n_samp, uint64_t seed, PhLin *out) {   /* :325-333 */
    PhaseSpace nom, *d = malloc(sizeof(PhaseSpace) * (n_samp > 0 ? n_samp : 1));
    ph_sample_target(t, NULL, &nom);
    McRng g = {seed};
    int k = 0;
    for (int i = 0; i < n_samp; i++) if (ph_sample_target(t, &g, &d[k]) == 0) k++;
    double sig[6][6];
    draw_cov(nom, d, k, sig);
    stm_track(nom, sig, tmax, out);
    free(d);
}

void ph_lin_cluster(const PhCluster *c, int n_samp, uint64_t seed, PhLin *out) {   /* phygas.c:335-343 */
    PhaseSpace nom, *d = malloc(sizeof(PhaseSpace) * (n_samp > 0 ? n_samp : 1));
    ph_sample_cluster(c, NULL, &nom);
    McRng g = {seed};
    int k = 0;
    for (int i = 0; i < n_samp; i++) if (ph_sample_cluster(c, &g, &d[k]) == 0) k++;
    double sig[6][6];
    draw_cov(nom, d, k, sig);
    stm_track(nom, sig, c->tmax_myr, out);
    free(d);
}

void ph_lin_free(PhLin *l) {
    free(l->pos); free(l->vel); free(l->S);
    l->pos = l->vel = NULL; l->S = NULL; l->n = 0;
}

/* ---- screening (reference: src/phygas.c:347-450) ---- */
static double mahal(const double M[6], Vec3 m) {   /* phygas.c:348-356 */
    double a = M[0], b = M[1], c = M[2], d = M[3], e = M[4], f = M[5];
    double c00 = d * f - e * e, c01 = c * e - b * f, c02 = b * e - c * d;
    double det = a * c00 + b * c01 + c * c02;
    if (!(det > 0)) return -1;
    double c11 = a * f - c * c, c12 = b * c - a * e, c22 = a * d - b * b;
    return (m.x * (c00 * m.x + c01 * m.y + c02 * m.z) + m.y * (c01 * m.x + c11 * m.y + c12 * m.z) +
            m.z * (c02 * m.x + c12 * m.y + c22 * m.z)) / det;
}

double orc_quad_inv(const double M[6], const double m[3]) { return mahal(M, mk(m[0], m[1], m[2])); }

double ph_screen_r = 2.0;

/* Scan with an exact cut (phygas.c:360-407); r_kernel as ph_screen_r. */
static double scan_pair(const PhLin *tl, const PhLin *cl, const PhCluster *c, double r_kernel, double chi2_max,
                        int *kb_out, double *sb_out) {
    int n = tl->n < cl->n ? tl->n : cl->n;
    double R = (r_kernel > 0 ? r_kernel * c->r50_pc : c->rtot_pc) / 1000.0, R2 = R * R;
    double best = INFINITY, sb = 0;
    int kb = 0;
    for (int k = 0; k < n; k++) {
        double M[6];
        for (int i = 0; i < 6; i++) M[i] = tl->S[k][i] + cl->S[k][i];
        M[0] += R2; M[3] += R2; M[5] += R2;
        Vec3 m0 = vsub(tl->pos[k], cl->pos[k]), m1 = m0;
        if (k + 1 < n) m1 = vsub(tl->pos[k + 1], cl->pos[k + 1]);
        double thr = 2.0 * chi2_max * (M[0] + M[3] + M[5]);
        if (thr < INFINITY) {
            Vec3 dr = vsub(m1, m0);
            double a = vdot(m0, dr), dd = vdot(dr, dr);
            int far;
            if (a >= 0) far = vdot(m0, m0) > thr;
            else if (-a >= dd) far = vdot(m1, m1) > thr;
            else far = vdot(m0, m0) * dd - a * a > thr * dd;
            if (far) continue;
        }
        double q = mahal(M, m0), s = 0;
        if (k + 1 < n) {
            segment_closest(m0, m1, &s);
            Vec3 ms = vadd(m0, vmul(vsub(m1, m0), s));
            double qs = mahal(M, ms);
            if (qs >= 0 && qs < q) q = qs; else s = 0;
        }
        if (q >= 0 && q < best) { best = q; kb = k; sb = s; }
    }
    *kb_out = kb;
    *sb_out = sb;
    return best;
}

double ph_screen_chi2_cut(const PhLin *tl, const PhLin *cl, const PhCluster *c, double chi2_max, int *kb,
                          double *sb) {
    return scan_pair(tl, cl, c, ph_screen_r, chi2_max, kb, sb);
}

double ph_screen_chi2(const PhLin *tl, const PhLin *cl, const PhCluster *c, int *kb, double *sb) {
    return scan_pair(tl, cl, c, ph_screen_r, INFINITY, kb, sb);
}

void ph_screen_finish(const PhLin *tl, const PhLin *cl, const PhCluster *c, uint64_t seed, double best, int kb,
                      double sb, PhScreen *res) {   /* phygas.c:413-444 */
    int n = tl->n < cl->n ? tl->n : cl->n;
    memset(res, 0, sizeof(*res));
    res->chi2 = best;
    res->t = (kb + sb) * PH_PRE_STEP_MYR;
    int k1 = kb + 1 < n ? kb + 1 : kb;
    Vec3 a = vsub(tl->pos[kb], cl->pos[kb]);
    Vec3 m = vadd(a, vmul(vsub(vsub(tl->pos[k1], cl->pos[k1]), a), sb));
    res->d_nom_pc = vlen(m) * 1000.0;
    res->vrel = vlen(vsub(tl->vel[kb], cl->vel[kb])) * KMS_PER_KPCMYR;
    double C[6];
    for (int i = 0; i < 6; i++) C[i] = tl->S[kb][i] + cl->S[kb][i];
    res->sig_pc = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
    double l00 = sqrt(fmax(C[0], 0)), l10 = l00 > 0 ? C[1] / l00 : 0, l20 = l00 > 0 ? C[2] / l00 : 0;
    double l11 = sqrt(fmax(C[3] - l10 * l10, 0)), l21 = l11 > 0 ? (C[4] - l20 * l10) / l11 : 0;
    double l22 = sqrt(fmax(C[5] - l20 * l20 - l21 * l21, 0));
    McRng g = {seed};
    double r50 = c->r50_pc / 1000.0, rt = c->rtot_pc / 1000.0;
    int in50 = 0, intot = 0;
    for (int i = 0; i < 512; i++) {
        double z0 = mc_gaussian(&g), z1 = mc_gaussian(&g), z2 = mc_gaussian(&g);
        Vec3 x = mk(m.x + l00 * z0, m.y + l10 * z0 + l11 * z1, m.z + l20 * z0 + l21 * z1 + l22 * z2);
        double r = vlen(x);
        if (r < r50) in50++;
        if (r < rt) intot++;
    }
    res->p_r50 = (double)in50 / 512;
    res->p_rtot = (double)intot / 512;
}

void ph_screen(const PhLin *tl, const PhLin *cl, const PhCluster *c, uint64_t seed, PhScreen *res) {
    int kb;
    double sb, best = ph_screen_chi2(tl, cl, c, &kb, &sb);
    ph_screen_finish(tl, cl, c, seed, best, kb, sb, res);
}

uint64_t ph_seed(const char *a, const char *b) {   /* phygas.c:452-458 */
    const uint64_t prime = 1099511628211ULL;
    uint64_t h = 1469598103934665603ULL;
    for (const unsigned char *p = (const unsigned char *)a; *p; p++) h = (h ^ *p) * prime;
    h = (h ^ (unsigned char)'|') * prime;
    for (const unsigned char *p = (const unsigned char *)b; *p; p++) h = (h ^ *p) * prime;
    return h;
}

/* ---- readers (reference: src/phygas.c:460-506) ---- */
static double num_or_nan(const char *s) { return strcmp(s, "nan") == 0 ? NAN : atof(s); }

int ph_read_clusters(const char *path, PhCluster **out) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int cap = 1024, n = 0, idx;
    PhCluster *v = malloc(sizeof(PhCluster) * cap);
    char line[1024], srv[32], serv[32], kind[4];
    while (fgets(line, sizeof line, f)) {
        if (n == cap) { cap *= 2; v = realloc(v, sizeof(PhCluster) * cap); }
        PhCluster *c = &v[n];
        if (sscanf(line, "%d %39s %3s %lf %lf %lf %lf %lf %lf %lf %lf %31s %31s %lf %lf %lf %lf %lf %lf", &idx,
                   c->name, kind, &c->ra, &c->dec, &c->dist_pc, &c->e_dist_pc, &c->pmra, &c->e_pmra, &c->pmdec,
                   &c->e_pmdec, srv, serv, &c->r50_pc, &c->rtot_pc, &c->age16, &c->age50, &c->age84,
                   &c->tmax_myr) != 19)
            continue;
        c->kind = kind[0];
        c->rv = num_or_nan(srv);
        c->e_rv = num_or_nan(serv);
        c->has_rv = !isnan(c->rv) && !isnan(c->e_rv);
        n++;
    }
    fclose(f);
    *out = v;
    return n;
}

int ph_read_targets(const char *path, PhTarget **out) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    int cap = 1024, n = 0;
    PhTarget *v = malloc(sizeof(PhTarget) * cap);
    char line[1024], srv[32], serv[32], kind[4];
    while (fgets(line, sizeof line, f)) {
        if (n == cap) { cap *= 2; v = realloc(v, sizeof(PhTarget) * cap); }
        PhTarget *t = &v[n];
        if (sscanf(line, "%39s %3s %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %31s %31s", t->id, kind, &t->ra, &t->dec,
                   &t->plx, &t->e_plx, &t->dist_pc, &t->e_lndist, &t->pmra, &t->e_pmra, &t->pmdec, &t->e_pmdec, srv,
                   serv) != 14)
            continue;
        t->kind = kind[0];
        t->rv = num_or_nan(srv);
        t->e_rv = num_or_nan(serv);
        t->has_rv = !isnan(t->rv) && !isnan(t->e_rv);
        n++;
    }
    fclose(f);
    *out = v;
    return n;
}

/* ======================= orc_* batch drivers (threads over independent items) ======================= */
typedef struct {
    void (*fn)(void *ctx, long i);
    void *ctx;
    long n;
    long next;
    pthread_mutex_t mu;
} Pool;

static void *pool_worker(void *arg) {
    Pool *p = arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        long i = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (i >= p->n) break;
        p->fn(p->ctx, i);
    }
    return NULL;
}

static void parallel_for(long n, int nthreads, void (*fn)(void *, long), void *ctx) {
    need_init();
    if (nthreads <= 1 || n <= 1) {
        for (long i = 0; i < n; i++) fn(ctx, i);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    Pool p = {fn, ctx, n, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, pool_worker, &p);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
}

typedef struct {
    const PhCluster *cl;
    const PhTarget *tg;
    int n_samp;
    double tmax;
    PhLin *lin;
} LinJob;

static void lin_cluster_job(void *ctx, long i) {
    LinJob *j = ctx;
    ph_lin_cluster(&j->cl[i], j->n_samp, ph_seed("lin", j->cl[i].name), &j->lin[i]);
}
static void lin_target_job(void *ctx, long i) {
    LinJob *j = ctx;
    ph_lin_target(&j->tg[i], j->tmax, j->n_samp, ph_seed("lin", j->tg[i].id), &j->lin[i]);
}

static long flatten(PhLin *lin, int cnt, int *n, long *off, double *pos, double *vel, double *S, long cap) {
    long tot = 0;
    for (int i = 0; i < cnt; i++) tot += lin[i].n;
    if (tot > cap) return -1;
    long o = 0;
    for (int i = 0; i < cnt; i++) {
        n[i] = lin[i].n;
        off[i] = o;
        for (int k = 0; k < lin[i].n; k++, o++) {
            pos[o * 3] = lin[i].pos[k].x; pos[o * 3 + 1] = lin[i].pos[k].y; pos[o * 3 + 2] = lin[i].pos[k].z;
            vel[o * 3] = lin[i].vel[k].x; vel[o * 3 + 1] = lin[i].vel[k].y; vel[o * 3 + 2] = lin[i].vel[k].z;
            for (int c = 0; c < 6; c++) S[o * 6 + c] = lin[i].S[k][c];
        }
    }
    return tot;
}

long orc_lin_clusters(const PhCluster *cl, int nc, int n_samp, int nthreads, int *n, long *off, double *pos,
                      double *vel, double *S, long cap) {
    PhLin *lin = calloc(nc > 0 ? nc : 1, sizeof(PhLin));
    LinJob j = {cl, NULL, n_samp, 0, lin};
    parallel_for(nc, nthreads, lin_cluster_job, &j);
    long tot = flatten(lin, nc, n, off, pos, vel, S, cap);
    for (int i = 0; i < nc; i++) ph_lin_free(&lin[i]);
    free(lin);
    return tot;
}

long orc_lin_targets(const PhTarget *tg, int nt, double tmax, int n_samp, int nthreads, int *n, long *off,
                     double *pos, double *vel, double *S, long cap) {
    PhLin *lin = calloc(nt > 0 ? nt : 1, sizeof(PhLin));
    LinJob j = {NULL, tg, n_samp, tmax, lin};
    parallel_for(nt, nthreads, lin_target_job, &j);
    long tot = flatten(lin, nt, n, off, pos, vel, S, cap);
    for (int i = 0; i < nt; i++) ph_lin_free(&lin[i]);
    free(lin);
    return tot;
}

void orc_lin_init_targets(const PhTarget *tg, int nt, int n_samp, double *ps0, double *sig) {
    for (int i = 0; i < nt; i++)
        ph_lin_target_init(&tg[i], n_samp, ph_seed("lin", tg[i].id), ps0 + (size_t)i * 6,
                           (double(*)[6])(sig + (size_t)i * 36));
}

typedef struct {
    const PhCluster *cl;
    const PhTarget *tg;
    int nc, t0, n_samp, use_cut;
    double tmax, chi2_max, r_kernel;
    PhLin *clin;
    OrcRow **rows;   /* per target */
    int *nrows;
    double *chi2;    /* orc_chi2_matrix */
    int *kb;
    double *sb;
} ScreenJob;

static void screen_target_job(void *ctx, long li) {
    ScreenJob *J = ctx;
    int j = J->t0 + (int)li;
    PhLin tl;
    ph_lin_target(&J->tg[j], J->tmax, J->n_samp, ph_seed("lin", J->tg[j].id), &tl);
    int cap = 16, cnt = 0;
    OrcRow *r = malloc(sizeof(OrcRow) * cap);
    for (int i = 0; i < J->nc; i++) {
        int kb;
        double sb, chi2 = scan_pair(&tl, &J->clin[i], &J->cl[i], J->r_kernel,
                                    J->use_cut ? J->chi2_max : INFINITY, &kb, &sb);
        if (!(chi2 < J->chi2_max)) continue;
        PhScreen s;
        ph_screen_finish(&tl, &J->clin[i], &J->cl[i], ph_seed(J->tg[j].id, J->cl[i].name), chi2, kb, sb, &s);
        if (cnt == cap) { cap *= 2; r = realloc(r, sizeof(OrcRow) * cap); }
        OrcRow *o = &r[cnt++];
        o->star = j; o->cluster = i; o->kb = kb; o->pad = 0;
        o->chi2 = s.chi2; o->sb = sb; o->t = s.t; o->d_nom_pc = s.d_nom_pc; o->sig_pc = s.sig_pc;
        o->p_r50 = s.p_r50; o->p_rtot = s.p_rtot; o->vrel = s.vrel;
    }
    ph_lin_free(&tl);
    J->rows[li] = r;
    J->nrows[li] = cnt;
}

static double max_tmax(const PhCluster *cl, int nc) {   /* phygas_screen.c:29-30 */
    double t = 0;
    for (int i = 0; i < nc; i++) if (cl[i].tmax_myr > t) t = cl[i].tmax_myr;
    return t;
}

long orc_screen_run(const PhCluster *cl, int nc, const PhTarget *tg, int nt, int t0, int t1, double chi2_max,
                    int n_samp, double r_kernel, int use_cut, int nthreads, OrcRow *rows, long cap) {
    if (t0 < 0) t0 = 0;
    if (t1 > nt) t1 = nt;
    int cnt = t1 > t0 ? t1 - t0 : 0;
    PhLin *clin = calloc(nc > 0 ? nc : 1, sizeof(PhLin));
    LinJob lj = {cl, NULL, n_samp, 0, clin};
    parallel_for(nc, nthreads, lin_cluster_job, &lj);
    ScreenJob J = {cl, tg, nc, t0, n_samp, use_cut, max_tmax(cl, nc), chi2_max, r_kernel, clin,
                   calloc(cnt > 0 ? cnt : 1, sizeof(OrcRow *)), calloc(cnt > 0 ? cnt : 1, sizeof(int)),
                   NULL, NULL, NULL};
    parallel_for(cnt, nthreads, screen_target_job, &J);
    long tot = 0;
    for (int i = 0; i < cnt; i++) tot += J.nrows[i];
    long o = 0;
    if (tot <= cap)
        for (int i = 0; i < cnt; i++)
            for (int k = 0; k < J.nrows[i]; k++) rows[o++] = J.rows[i][k];
    for (int i = 0; i < cnt; i++) free(J.rows[i]);
    free(J.rows); free(J.nrows);
    for (int i = 0; i < nc; i++) ph_lin_free(&clin[i]);
    free(clin);
    return tot <= cap ? tot : -tot;
}

static void chi2_target_job(void *ctx, long j) {
    ScreenJob *J = ctx;
    PhLin tl;
    ph_lin_target(&J->tg[j], J->tmax, J->n_samp, ph_seed("lin", J->tg[j].id), &tl);
    for (int i = 0; i < J->nc; i++) {
        size_t id = (size_t)j * J->nc + i;
        J->chi2[id] = scan_pair(&tl, &J->clin[i], &J->cl[i], J->r_kernel, INFINITY, &J->kb[id], &J->sb[id]);
    }
    ph_lin_free(&tl);
}

void orc_chi2_matrix(const PhCluster *cl, int nc, const PhTarget *tg, int nt, int n_samp, double r_kernel,
                     int nthreads, double *chi2, int *kb, double *sb) {
    PhLin *clin = calloc(nc > 0 ? nc : 1, sizeof(PhLin));
    LinJob lj = {cl, NULL, n_samp, 0, clin};
    parallel_for(nc, nthreads, lin_cluster_job, &lj);
    ScreenJob J = {cl, tg, nc, 0, n_samp, 0, max_tmax(cl, nc), INFINITY, r_kernel, clin, NULL, NULL, chi2, kb, sb};
    parallel_for(nt, nthreads, chi2_target_job, &J);
    for (int i = 0; i < nc; i++) ph_lin_free(&clin[i]);
    free(clin);
}

typedef struct {
    const PhTarget *tg;
    const PhCluster *cl;
    const int *ti, *ci;
    int n_clones;
    double dt_myr;
    PhResult *res;
    const double *ps;
    const int *steps;
    const double *dts;
    double *best, *tbest, *vbest;
} PairJob;

static void pair_job(void *ctx, long i) {
    PairJob *J = ctx;
    const PhTarget *t = &J->tg[J->ti[i]];
    const PhCluster *c = &J->cl[J->ci[i]];
    ph_pair(t, c, J->n_clones, J->dt_myr, ph_seed(t->id, c->name), &J->res[i]);
}

void orc_pair_run(const PhTarget *tg, const PhCluster *cl, int npairs, const int *ti, const int *ci, int n_clones,
                  double dt_myr, int nthreads, PhResult *res) {
    PairJob J = {tg, cl, ti, ci, n_clones, dt_myr, res, NULL, NULL, NULL, NULL, NULL, NULL};
    parallel_for(npairs, nthreads, pair_job, &J);
}

static void clone_job(void *ctx, long i) {
    PairJob *J = ctx;
    ph_pair_clone(J->ps + (size_t)i * 12, J->steps[i], J->dts[i], &J->best[i], &J->tbest[i], &J->vbest[i]);
}

void orc_pair_clones(const double *ps, const int *steps, const double *dts, int n, int nthreads, double *best,
                     double *tbest, double *vbest) {
    PairJob J = {NULL, NULL, NULL, NULL, 0, 0, NULL, ps, steps, dts, best, tbest, vbest};
    parallel_for(n, nthreads, clone_job, &J);
}

/* Pair-mode prefilter over every (target, cluster), as phygas_cli runs it (reference:
 * src/phygas_cli.c:103-118): cluster tracks with 32 spread draws seeded "track"|name, target tracks
 * over the longest cluster tmax seeded "track"|id.  mask[j * nc + i] = ph_prefilter(...). */
void orc_prefilter_mask(const PhTarget *tg, int nt, const PhCluster *cl, int nc, unsigned char *mask) {
    need_init();
    double tmax = max_tmax(cl, nc);
    PhTrack *ct = malloc(sizeof(PhTrack) * (nc > 0 ? nc : 1));
    for (int i = 0; i < nc; i++) ph_track_cluster(&cl[i], 32, ph_seed("track", cl[i].name), &ct[i]);
    for (int j = 0; j < nt; j++) {
        PhTrack tt;
        ph_track_target(&tg[j], tmax, 32, ph_seed("track", tg[j].id), &tt);
        for (int i = 0; i < nc; i++) mask[(size_t)j * nc + i] = (unsigned char)ph_prefilter(&tt, &ct[i], &cl[i]);
        ph_track_free(&tt);
    }
    for (int i = 0; i < nc; i++) ph_track_free(&ct[i]);
    free(ct);
}

#ifdef ORC_ULP_SHIM
#undef pow
#undef sin
#undef cos
#undef hypot
static unsigned g_ulp_mask;
static double g_ulp_rate;

/* mask bits: 1 pow, 2 sin, 4 cos, 8 hypot; rate: the share of calls moved */
void orc_ulp_set(unsigned mask, double rate) {
    g_ulp_mask = mask;
    g_ulp_rate = rate;
}

static double ulp_move(unsigned bit, double v, double x, double y) {
    if (!(g_ulp_mask & bit)) return v;
    uint64_t bx, by;
    memcpy(&bx, &x, 8);
    memcpy(&by, &y, 8);
    uint64_t h = bx * 0x9E3779B97F4A7C15ULL ^ (by + bit) * 0xBF58476D1CE4E5B9ULL;
    h = (h ^ (h >> 31)) * 0x94D049BB133111EBULL;
    h ^= h >> 29;
    if ((double)(h >> 11) * (1.0 / 9007199254740992.0) < g_ulp_rate) v = nextafter(v, (h & 1) ? INFINITY : -INFINITY);
    return v;
}
double orc_ulp_pow(double x, double y) { return ulp_move(1, pow(x, y), x, y); }
double orc_ulp_sin(double x) { return ulp_move(2, sin(x), x, 0.0); }
double orc_ulp_cos(double x) { return ulp_move(4, cos(x), x, 0.0); }
double orc_ulp_hypot(double x, double y) { return ulp_move(8, hypot(x, y), x, y); }
#endif
