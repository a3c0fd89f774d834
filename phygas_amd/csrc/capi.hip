// The reference's core C API (reference: src/phygas.h:64-139, src/montecarlo.h:48, src/potential.h:13-19) on the
// library's host physics: physics.h compiled for the host calls the C library's pow / sin / cos / hypot / log /
// exp as the reference does and the library is built without FMA contraction, so every result here is the
// reference's, bit for bit.  These are the reference's synchronous per-object calls, for programs written
// against it (its CLIs link this library in place of phygas.o); batches of stars, clusters and clones run on
// the device through the phx_* engine with the same results.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <type_traits>
#include <vector>

#include "common.h"
#include "phygas_core.h"

using namespace phx;

namespace {

const Pot &pot() {
    static const Pot P = default_pot();
    return P;
}

Vec3 vec(V3 a) { Vec3 r = {a.x, a.y, a.z}; return r; }
V3 v3of(Vec3 a) { return v3(a.x, a.y, a.z); }
void to_ps(V3 p, V3 v, PhaseSpace *o) {
    o->pos = vec(p);
    o->vel = vec(v);
}

void pot_accel(const Potential *, Vec3 pos, Vec3 *out) { *out = vec(accel(pot(), v3of(pos))); }

// Potential energy (reference: src/potential.c:103-136): Plummer bulge, Miyamoto-Nagai disk, and the halo's
// potential from its enclosed mass, the exterior of the cutoff a point mass.
double halo_gm(const Pot &P, double r) {
    const double re = r < P.halo_cut ? r : P.halo_cut, x = re / P.halo_a;
    return P.halo_gm * pow(x, P.halo_gamma) / (1.0 + pow(x, P.halo_gamma - 1.0));
}
double pot_energy(const Potential *, Vec3 pos) {
    const Pot &P = pot();
    const double R2 = pos.x * pos.x + pos.y * pos.y;
    const double r = norm(v3of(pos));
    const double phb = -P.bulge_gm / sqrt(r * r + P.bulge_b * P.bulge_b);
    const double zeta = sqrt(pos.z * pos.z + P.disk_b * P.disk_b), az = P.disk_a + zeta;
    const double phd = -P.disk_gm / sqrt(R2 + az * az);
    const double rs = r > 1e-12 ? r : 1e-12;
    const double gm_cut = halo_gm(P, P.halo_cut);
    double phh;
    if (rs >= P.halo_cut) {
        phh = -gm_cut / rs;
    } else {
        const double g1 = P.halo_gamma - 1.0, xc = P.halo_cut / P.halo_a, xr = rs / P.halo_a;
        const double integ = (1.0 / g1) * (log(1.0 + pow(xc, g1)) - log(1.0 + pow(xr, g1)));
        phh = -(P.halo_gm / P.halo_a) * integ - gm_cut / P.halo_cut;
    }
    return phb + phd + phh;
}

const Potential g_potential = {pot_accel, pot_energy, nullptr};

// Nominal draw and up to cap seeded draws (failures skipped) of a record.
template <class Rec>
int draws(const Rec &r, uint64_t seed, int n, int cap, V3 &np, V3 &nv, std::vector<V3> &dp, std::vector<V3> &dv) {
    constexpr bool cluster = std::is_same<Rec, PhCluster>::value;
    np = nv = v3(0, 0, 0);
    if (cluster) sample_cluster(pot(), (const PhCluster &)r, nullptr, np, nv);
    else sample_target(pot(), (const PhTarget &)r, nullptr, np, nv);
    if (n > cap) n = cap;
    uint64_t st = seed;
    dp.clear();
    dv.clear();
    for (int i = 0; i < n; i++) {
        V3 p, v;
        const int rc = cluster ? sample_cluster(pot(), (const PhCluster &)r, &st, p, v)
                               : sample_target(pot(), (const PhTarget &)r, &st, p, v);
        if (rc == 0) {
            dp.push_back(p);
            dv.push_back(v);
        }
    }
    return (int)dp.size();
}

// Nominal path at 0.5-Myr points, 10 kick-drift-kick steps each (reference: src/phygas.c:98-110), and the RMS
// spread of the draws about the nominal (:112-120).
void path(V3 np, V3 nv, double tmax, const std::vector<V3> &dp, const std::vector<V3> &dv, PhTrack *out) {
    const int n = (int)ceil(tmax / STEP_MYR) + 1;
    out->n = n;
    out->pos = (Vec3 *)malloc(sizeof(Vec3) * (size_t)n);
    Body b = body_start(pot(), np, nv);
    const double dt = -STEP_MYR / 10;
    out->pos[0] = vec(b.pos);
    for (int k = 1; k < n; k++) {
        for (int i = 0; i < 10; i++) kdk(pot(), b, dt);
        out->pos[k] = vec(b.pos);
    }
    double sp = 0, sv = 0;
    for (size_t i = 0; i < dp.size(); i++) {
        const V3 a = dp[i] - np, c = dv[i] - nv;
        sp += dot(a, a);
        sv += dot(c, c);
    }
    const int k = (int)dp.size();
    out->sig_pos = k ? sqrt(sp / k) : 0;
    out->sig_vel = k ? sqrt(sv / k) : 0;
}

// 6x6 covariance of the draws about the nominal, divided by their count (reference: src/phygas.c:252-264)
void lin_sigma(V3 np, V3 nv, const std::vector<V3> &dp, const std::vector<V3> &dv, double sig[6][6]) {
    memset(sig, 0, sizeof(double[6][6]));
    const int nd = (int)dp.size();
    for (int i = 0; i < nd; i++) {
        const double d[6] = {dp[i].x - np.x, dp[i].y - np.y, dp[i].z - np.z, dv[i].x - nv.x, dv[i].y - nv.y, dv[i].z - nv.z};
        for (int a = 0; a < 6; a++)
            for (int b = 0; b < 6; b++) sig[a][b] += d[a] * d[b];
    }
    for (int a = 0; a < 6; a++)
        for (int b = 0; b < 6; b++) sig[a][b] /= nd > 0 ? nd : 1;
}

// Nominal orbit and its six neighbours (offsets 1e-5 kpc / 1e-5 kpc/Myr), the finite-difference position rows A
// of the state-transition matrix and S = A Sigma0 A^T at every 0.5-Myr point (reference: src/phygas.c:286-323).
void lin_tracks(V3 np, V3 nv, const double sig[6][6], double tmax, PhLin *out) {
    Body bd[7];
    bd[0] = body_start(pot(), np, nv);
    for (int j = 0; j < 6; j++) {
        double c[6] = {np.x, np.y, np.z, nv.x, nv.y, nv.z};
        c[j] += j < 3 ? PH_LIN_EPS_POS : PH_LIN_EPS_VEL;
        bd[j + 1] = body_start(pot(), v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]));
    }
    const int n = (int)ceil(tmax / STEP_MYR) + 1;
    out->n = n;
    out->pos = (Vec3 *)malloc(sizeof(Vec3) * (size_t)n);
    out->vel = (Vec3 *)malloc(sizeof(Vec3) * (size_t)n);
    out->S = (double(*)[6])malloc(sizeof(double[6]) * (size_t)n);
    const double dt = -STEP_MYR / PH_LIN_SUB;
    for (int k = 0; k < n; k++) {
        if (k > 0)
            for (int i = 0; i < PH_LIN_SUB; i++)
                for (int j = 0; j < 7; j++) kdk(pot(), bd[j], dt);
        out->pos[k] = vec(bd[0].pos);
        out->vel[k] = vec(bd[0].vel);
        double A[3][6];
        for (int j = 0; j < 6; j++) {
            const double e = j < 3 ? PH_LIN_EPS_POS : PH_LIN_EPS_VEL;
            const V3 d = bd[j + 1].pos - bd[0].pos;
            A[0][j] = d.x / e;
            A[1][j] = d.y / e;
            A[2][j] = d.z / e;
        }
        double AS[3][6] = {{0}}, S[3][3] = {{0}};
        for (int r = 0; r < 3; r++)
            for (int b = 0; b < 6; b++)
                for (int a = 0; a < 6; a++) AS[r][b] += A[r][a] * sig[a][b];
        for (int r = 0; r < 3; r++)
            for (int s = 0; s < 3; s++)
                for (int b = 0; b < 6; b++) S[r][s] += AS[r][b] * A[s][b];
        out->S[k][0] = S[0][0]; out->S[k][1] = S[0][1]; out->S[k][2] = S[0][2];
        out->S[k][3] = S[1][1]; out->S[k][4] = S[1][2]; out->S[k][5] = S[2][2];
    }
}

V3 lin_pos(const PhLin *l, int k) { return v3of(l->pos[k]); }

int cmp_double(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

// linear interpolation at q (n - 1) of a sorted column (reference: src/phygas.c:175-180)
double quantile(const double *v, int n, double q) {
    const double pos = q * (n - 1);
    const int i = (int)pos;
    if (i >= n - 1) return v[n - 1];
    return v[i] + (pos - i) * (v[i + 1] - v[i]);
}

template <class Rec>
int read_records(const char *path, Rec **out) {
    std::vector<std::string> lines;
    if (!read_lines(path, lines)) return -1;
    std::vector<Rec> v;
    for (const std::string &s : lines) {
        Rec r;
        bool ok;
        if constexpr (std::is_same<Rec, PhCluster>::value) {
            ok = parse_cluster(s.c_str(), (PhCluster *)&r);
        } else {
            std::string only;
            ok = parse_target(s.c_str(), (PhTarget *)&r, &only);
        }
        if (ok) v.push_back(r);
    }
    Rec *o = (Rec *)malloc(sizeof(Rec) * (v.empty() ? 1 : v.size()));
    if (!o) return -1;
    if (!v.empty()) memcpy(o, v.data(), sizeof(Rec) * v.size());
    *out = o;
    return (int)v.size();
}

}  // namespace

extern "C" {

double ph_screen_r = 2.0;

double mc_gaussian(McRng *rng) { return gaussian(rng->state); }

void ph_init(void) {}

const Potential *ph_potential(void) { return &g_potential; }

int ph_sample_target(const PhTarget *t, McRng *rng, PhaseSpace *out) {
    V3 p = v3(0, 0, 0), v = v3(0, 0, 0);
    const int rc = sample_target(pot(), *t, rng ? &rng->state : nullptr, p, v);
    if (rc == 0) to_ps(p, v, out);
    return rc;
}

int ph_sample_cluster(const PhCluster *c, McRng *rng, PhaseSpace *out) {
    V3 p = v3(0, 0, 0), v = v3(0, 0, 0);
    const int rc = sample_cluster(pot(), *c, rng ? &rng->state : nullptr, p, v);
    if (rc == 0) to_ps(p, v, out);
    return rc;
}

void ph_track_target(const PhTarget *t, double tmax, int n_spread, uint64_t seed, PhTrack *out) {
    V3 np, nv;
    std::vector<V3> dp, dv;
    draws(*t, seed, n_spread, 64, np, nv, dp, dv);
    path(np, nv, tmax, dp, dv, out);
}

void ph_track_cluster(const PhCluster *c, int n_spread, uint64_t seed, PhTrack *out) {
    V3 np, nv;
    std::vector<V3> dp, dv;
    draws(*c, seed, n_spread, 64, np, nv, dp, dv);
    path(np, nv, c->tmax_myr, dp, dv, out);
}

void ph_track_free(PhTrack *t) {
    free(t->pos);
    t->pos = nullptr;
    t->n = 0;
}

// reference: src/phygas.c:158-168
int ph_prefilter(const PhTrack *tt, const PhTrack *ct, const PhCluster *c) {
    const int n = tt->n < ct->n ? tt->n : ct->n;
    const double sp = hypot(tt->sig_pos, ct->sig_pos), sv = hypot(tt->sig_vel, ct->sig_vel);
    const double r0 = (3.0 * c->rtot_pc + 10.0) / 1000.0;
    for (int k = 0; k + 1 < n; k++) {
        const V3 m0 = v3of(tt->pos[k]) - v3of(ct->pos[k]), m1 = v3of(tt->pos[k + 1]) - v3of(ct->pos[k + 1]);
        const double s = seg_param(m0, m1);
        const double d = norm(m0 + (m1 - m0) * s);
        if (d < r0 + 3.0 * (sp + sv * ((k + s) * STEP_MYR))) return 1;
    }
    return 0;
}

int ph_pair_draw(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, double *ps,
                 int *steps, double *dt) {
    return phx_pair_draw(t, c, n_clones, dt_myr, seed, ps, steps, dt);
}

// reference: src/phygas.c:200-216
void ph_pair_clone(const double *ps, int steps, double dt, double *best_out, double *tbest_out, double *vbest_out) {
    Body a = body_start(pot(), v3(ps[0], ps[1], ps[2]), v3(ps[3], ps[4], ps[5]));
    Body b = body_start(pot(), v3(ps[6], ps[7], ps[8]), v3(ps[9], ps[10], ps[11]));
    V3 r0 = a.pos - b.pos;
    double best = norm(r0), tbest = 0, vbest = norm(a.vel - b.vel);
    for (int i = 1; i <= steps; i++) {
        kdk(pot(), a, dt);
        kdk(pot(), b, dt);
        const V3 r1 = a.pos - b.pos;
        const double s = seg_param(r0, r1);
        const double dm = norm(r0 + (r1 - r0) * s);
        if (dm < best) {
            best = dm;
            tbest = (i - 1 + s) * -dt;
            vbest = norm(a.vel - b.vel);
        }
        r0 = r1;
    }
    *best_out = best;
    *tbest_out = tbest;
    *vbest_out = vbest;
}

// reference: src/phygas.c:218-238
void ph_pair_reduce(const PhCluster *c, int m, const double *best, const double *tbest, const double *vbest,
                    PhResult *res) {
    std::vector<double> d((size_t)(m > 0 ? m : 1)), tt(d.size()), vv(d.size());
    int in50 = 0, intot = 0;
    for (int k = 0; k < m; k++) {
        d[k] = best[k] * 1000.0;
        tt[k] = tbest[k];
        vv[k] = vbest[k] * KMS_PER_KPCMYR;
        if (d[k] < c->r50_pc) in50++;
        if (d[k] < c->rtot_pc) intot++;
    }
    memset(res, 0, sizeof *res);
    res->n = m;
    if (m <= 0) return;
    res->f_r50 = (double)in50 / m;
    res->f_rtot = (double)intot / m;
    qsort(d.data(), (size_t)m, sizeof(double), cmp_double);
    qsort(tt.data(), (size_t)m, sizeof(double), cmp_double);
    qsort(vv.data(), (size_t)m, sizeof(double), cmp_double);
    res->d05 = quantile(d.data(), m, 0.05); res->d50 = quantile(d.data(), m, 0.5); res->d95 = quantile(d.data(), m, 0.95);
    res->t05 = quantile(tt.data(), m, 0.05); res->t50 = quantile(tt.data(), m, 0.5); res->t95 = quantile(tt.data(), m, 0.95);
    res->v05 = quantile(vv.data(), m, 0.05); res->v50 = quantile(vv.data(), m, 0.5); res->v95 = quantile(vv.data(), m, 0.95);
}

void ph_pair(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, PhResult *res) {
    const size_t cap = (size_t)(n_clones > 0 ? n_clones : 1);
    std::vector<double> ps(12 * cap), out(3 * cap);
    int steps;
    double dt;
    const int m = ph_pair_draw(t, c, n_clones, dt_myr, seed, ps.data(), &steps, &dt);
    for (int k = 0; k < m; k++)
        ph_pair_clone(ps.data() + (size_t)k * 12, steps, dt, &out[k], &out[m + k], &out[2 * (size_t)m + k]);
    ph_pair_reduce(c, m, out.data(), out.data() + m, out.data() + 2 * (size_t)m, res);
}

// reference: src/phygas.c:274-284
void ph_lin_target_init(const PhTarget *t, int n_samp, uint64_t seed, double ps0[6], double sig[6][6]) {
    V3 np, nv;
    std::vector<V3> dp, dv;
    draws(*t, seed, n_samp, n_samp, np, nv, dp, dv);
    lin_sigma(np, nv, dp, dv, sig);
    ps0[0] = np.x; ps0[1] = np.y; ps0[2] = np.z; ps0[3] = nv.x; ps0[4] = nv.y; ps0[5] = nv.z;
}

// reference: src/phygas.c:325-333
void ph_lin_target(const PhTarget *t, double tmax, int n_samp, uint64_t seed, PhLin *out) {
    V3 np, nv;
    std::vector<V3> dp, dv;
    double sig[6][6];
    draws(*t, seed, n_samp, n_samp, np, nv, dp, dv);
    lin_sigma(np, nv, dp, dv, sig);
    lin_tracks(np, nv, sig, tmax, out);
}

// reference: src/phygas.c:335-343
void ph_lin_cluster(const PhCluster *c, int n_samp, uint64_t seed, PhLin *out) {
    V3 np, nv;
    std::vector<V3> dp, dv;
    double sig[6][6];
    draws(*c, seed, n_samp, n_samp, np, nv, dp, dv);
    lin_sigma(np, nv, dp, dv, sig);
    lin_tracks(np, nv, sig, c->tmax_myr, out);
}

void ph_lin_free(PhLin *l) {
    free(l->pos);
    free(l->vel);
    free(l->S);
    l->pos = l->vel = nullptr;
    l->S = nullptr;
    l->n = 0;
}

// The chi2 scan over time with an exact early reject (reference: src/phygas.c:358-407): a point whose segment
// stays farther from the origin than sqrt(2 chi2_max trace(M)) cannot bring chi2 below the cut (m^T M^-1 m >=
// |m|^2 / trace(M)), so it is skipped; pairs below the cut get the same chi2, kb and sb as the full scan.
double ph_screen_chi2_cut(const PhLin *tl, const PhLin *cl, const PhCluster *c, double chi2_max, int *kb_out,
                          double *sb_out) {
    const int n = tl->n < cl->n ? tl->n : cl->n;
    const double R = (ph_screen_r > 0 ? ph_screen_r * c->r50_pc : c->rtot_pc) / 1000.0, R2 = R * R;
    double best = INFINITY, sb = 0;
    int kb = 0;
    for (int k = 0; k < n; k++) {
        double M[6];
        for (int i = 0; i < 6; i++) M[i] = tl->S[k][i] + cl->S[k][i];
        M[0] += R2; M[3] += R2; M[5] += R2;
        const V3 m0 = lin_pos(tl, k) - lin_pos(cl, k);
        const V3 m1 = k + 1 < n ? lin_pos(tl, k + 1) - lin_pos(cl, k + 1) : m0;
        const double thr = 2.0 * chi2_max * (M[0] + M[3] + M[5]);
        if (thr < INFINITY) {
            const V3 dr = m1 - m0;
            const double a = dot(m0, dr), dd = dot(dr, dr);
            bool far;
            if (a >= 0) far = dot(m0, m0) > thr;
            else if (-a >= dd) far = dot(m1, m1) > thr;
            else far = dot(m0, m0) * dd - a * a > thr * dd;
            if (far) continue;
        }
        double q = mahalanobis(M, m0.x, m0.y, m0.z), s = 0;
        if (k + 1 < n) {
            s = seg_param(m0, m1);
            const V3 ms = m0 + (m1 - m0) * s;
            const double qs = mahalanobis(M, ms.x, ms.y, ms.z);
            if (qs >= 0 && qs < q) q = qs; else s = 0;
        }
        if (q >= 0 && q < best) { best = q; kb = k; sb = s; }
    }
    *kb_out = kb;
    *sb_out = sb;
    return best;
}

double ph_screen_chi2(const PhLin *tl, const PhLin *cl, const PhCluster *c, int *kb, double *sb) {
    return ph_screen_chi2_cut(tl, cl, c, INFINITY, kb, sb);
}

// reference: src/phygas.c:413-444
void ph_screen_finish(const PhLin *tl, const PhLin *cl, const PhCluster *c, uint64_t seed, double best, int kb,
                      double sb, PhScreen *res) {
    const int n = tl->n < cl->n ? tl->n : cl->n;
    memset(res, 0, sizeof *res);
    res->chi2 = best;
    res->t = (kb + sb) * STEP_MYR;
    const int k1 = kb + 1 < n ? kb + 1 : kb;
    const V3 a = lin_pos(tl, kb) - lin_pos(cl, kb), b = lin_pos(tl, k1) - lin_pos(cl, k1);
    const V3 m = a + (b - a) * sb;
    res->d_nom_pc = norm(m) * 1000.0;
    res->vrel = norm(v3of(tl->vel[kb]) - v3of(cl->vel[kb])) * KMS_PER_KPCMYR;
    double C[6];
    for (int i = 0; i < 6; i++) C[i] = tl->S[kb][i] + cl->S[kb][i];
    res->sig_pc = sqrt((C[0] + C[3] + C[5]) / 3.0) * 1000.0;
    int in50, intot;
    finish_counts(C, m, seed, c->r50_pc / 1000.0, c->rtot_pc / 1000.0, in50, intot);
    res->p_r50 = (double)in50 / 512;
    res->p_rtot = (double)intot / 512;
}

void ph_screen(const PhLin *tl, const PhLin *cl, const PhCluster *c, uint64_t seed, PhScreen *res) {
    int kb;
    double sb;
    const double best = ph_screen_chi2(tl, cl, c, &kb, &sb);
    ph_screen_finish(tl, cl, c, seed, best, kb, sb, res);
}

uint64_t ph_seed(const char *a, const char *b) { return phx_seed(a, b); }

int ph_read_clusters(const char *path, PhCluster **out) { return read_records(path, out); }
int ph_read_targets(const char *path, PhTarget **out) { return read_records(path, out); }

}  // extern "C"
