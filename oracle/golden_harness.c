/* Golden-vector harness -- test infrastructure.
 *
 * One source, built twice by oracle/Makefile: against the compiled reference (-DUSE_REF, headers from
 * the reference's src/) into oracle/_ref/golden_ref, and against the oracle (oracle/liboracle.so) into
 * oracle/golden_orc.  Both print the same records as hex floats; tests/golden/ holds the reference's
 * output and tests/test_oracle.py requires the oracle's to be identical, byte for byte.  tests/test_capi.py
 * builds it a third time (-DUSE_CORE) against the product's core API (include/phygas_core.h,
 * phygas_amd/libphygas_hip.so), whose output must be identical too (every section but coords, which is not
 * part of that API).
 *
 *   golden_harness <section> clusters.txt targets.txt
 * sections: accel rng seed coords sample lininit lin chi2 pair track read */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef USE_REF
#include "phygas.h"
static void to_gc(double ra, double dec, double plx, double pmra, double pmdec, double rv, double o[6]) {
    Astrometry a = {ra, dec, plx, pmra, pmdec, rv};
    PhaseSpace p;
    astrometry_to_galactocentric(a, &solar_params_default, &p);
    o[0] = p.pos.x; o[1] = p.pos.y; o[2] = p.pos.z; o[3] = p.vel.x; o[4] = p.vel.y; o[5] = p.vel.z;
}
static void from_gc(const double ps[6], double o[6]) {
    PhaseSpace p = {{ps[0], ps[1], ps[2]}, {ps[3], ps[4], ps[5]}};
    Astrometry a = galactocentric_to_astrometry(p, &solar_params_default);
    o[0] = a.ra_deg; o[1] = a.dec_deg; o[2] = a.parallax_mas; o[3] = a.pmra_masyr; o[4] = a.pmdec_masyr;
    o[5] = a.rv_kms;
}
#elif defined(USE_CORE)
#include "phygas_core.h"
static void to_gc(double ra, double dec, double plx, double pmra, double pmdec, double rv, double o[6]) {
    (void)ra; (void)dec; (void)plx; (void)pmra; (void)pmdec; (void)rv; (void)o;
    fprintf(stderr, "coords: not part of the core API\n");
    exit(2);
}
static void from_gc(const double ps[6], double o[6]) { (void)ps; (void)o; exit(2); }
#else
#include "phygas_oracle.h"
static void to_gc(double ra, double dec, double plx, double pmra, double pmdec, double rv, double o[6]) {
    orc_astrometry_to_gc(ra, dec, plx, pmra, pmdec, rv, o);
}
static void from_gc(const double ps[6], double o[6]) { orc_gc_to_astrometry(ps, o); }
#endif

static void pv(const char *tag, int i, const double *v, int n) {
    printf("%s %d", tag, i);
    for (int k = 0; k < n; k++) printf(" %a", v[k]);
    printf("\n");
}

static double max_tmax(const PhCluster *cl, int nc) {
    double t = 0;
    for (int i = 0; i < nc; i++) if (cl[i].tmax_myr > t) t = cl[i].tmax_myr;
    return t;
}

static void print_lin(const char *tag, int i, const PhLin *l) {
    printf("%s_n %d %d\n", tag, i, l->n);
    for (int k = 0; k < l->n; k++) {
        if (k % 10 != 0 && k + 1 != l->n && k > 3) continue;
        double v[13] = {k, l->pos[k].x, l->pos[k].y, l->pos[k].z, l->vel[k].x, l->vel[k].y, l->vel[k].z,
                        l->S[k][0], l->S[k][1], l->S[k][2], l->S[k][3], l->S[k][4], l->S[k][5]};
        pv(tag, i, v, 13);
    }
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s section clusters.txt targets.txt\n", argv[0]);
        return 1;
    }
    const char *sec = argv[1];
    ph_init();
    PhCluster *cl;
    PhTarget *tg;
    int nc = ph_read_clusters(argv[2], &cl), nt = ph_read_targets(argv[3], &tg);
    if (nc <= 0 || nt <= 0) { fprintf(stderr, "bad inputs %d %d\n", nc, nt); return 1; }

    if (!strcmp(sec, "accel")) {
        static const double pts[][3] = {
            {-8.122, 0, 0.0208}, {0, 0, 0}, {0, 0, 1}, {1e-13, 0, 0}, {0, 0, -0.5}, {3, 4, 0.5}, {-3, -4, -0.5},
            {150, 0, 0}, {99.99, 0, 0}, {100, 0, 0}, {100.01, 0, 0}, {0.2, -0.1, 0.05}, {-7.5, 1.2, -0.3},
            {-9.9, -2.2, 0.8}, {12.0, 12.0, 12.0}, {70, 70, 10}, {-8.0, 0.5, 1e-9}, {1e-12, 1e-12, 0}};
        const Potential *pot = ph_potential();
        for (int i = 0; i < (int)(sizeof pts / sizeof pts[0]); i++) {
            Vec3 p = {pts[i][0], pts[i][1], pts[i][2]}, a;
            pot->acceleration(pot, p, &a);
            double v[4] = {a.x, a.y, a.z, pot->energy(pot, p)};
            pv("accel", i, v, 4);
        }
        /* a dense set of points along tracks to exercise every branch of pow */
        for (int i = 0; i < 400; i++) {
            double t = i * 0.05;
            Vec3 p = {-8.0 + 3.0 * sin(t), 2.0 * cos(1.3 * t), 0.4 * sin(2.1 * t)}, a;
            pot->acceleration(pot, p, &a);
            double v[3] = {a.x, a.y, a.z};
            pv("accel_line", i, v, 3);
        }
    } else if (!strcmp(sec, "rng")) {
        const unsigned long long seeds[] = {0ULL, 1ULL, 42ULL, 0xFFFFFFFFFFFFFFFFULL, 0x9E3779B97F4A7C15ULL,
                                            (unsigned long long)ph_seed("lin", "C00001")};
        for (int i = 0; i < 6; i++) {
            McRng g;
            mc_rng_seed(&g, seeds[i]);
            for (int k = 0; k < 8; k++) printf("u64 %d %llu\n", i, (unsigned long long)mc_rng_next_u64(&g));
            mc_rng_seed(&g, seeds[i]);
            double v[32];
            for (int k = 0; k < 32; k++) v[k] = mc_gaussian(&g);
            pv("gauss", i, v, 32);
            mc_rng_seed(&g, seeds[i]);
            for (int k = 0; k < 8; k++) v[k] = mc_rng_uniform(&g);
            pv("unif", i, v, 8);
        }
    } else if (!strcmp(sec, "seed")) {
        const char *a[] = {"", "lin", "track", "S0000001", "J0534+2200", "\xff\x80x", "a|b"};
        const char *b[] = {"", "C00001", "Alessi_12", "UBC 1 (m)", "zz", "|"};
        for (int i = 0; i < 7; i++)
            for (int j = 0; j < 6; j++) printf("seed %d %d %llu\n", i, j, (unsigned long long)ph_seed(a[i], b[j]));
        for (int i = 0; i < nt; i++) printf("seed_t %d %llu\n", i, (unsigned long long)ph_seed("lin", tg[i].id));
    } else if (!strcmp(sec, "coords")) {
        for (int i = 0; i < nt; i++) {
            double o[6], b[6];
            double plx = tg[i].plx > 0 ? tg[i].plx : 1000.0 / tg[i].dist_pc;
            to_gc(tg[i].ra, tg[i].dec, plx, tg[i].pmra, tg[i].pmdec, tg[i].has_rv ? tg[i].rv : 0.0, o);
            pv("to_gc", i, o, 6);
            from_gc(o, b);
            pv("from_gc", i, b, 6);
        }
        for (int i = 0; i < nc; i++) {
            double o[6];
            to_gc(cl[i].ra, cl[i].dec, 1000.0 / cl[i].dist_pc, cl[i].pmra, cl[i].pmdec, cl[i].has_rv ? cl[i].rv : 0.0,
                  o);
            pv("to_gc_c", i, o, 6);
        }
    } else if (!strcmp(sec, "sample")) {
        for (int i = 0; i < nt; i++) {
            PhaseSpace p;
            int rc = ph_sample_target(&tg[i], NULL, &p);
            double v[7] = {rc, p.pos.x, p.pos.y, p.pos.z, p.vel.x, p.vel.y, p.vel.z};
            pv("tnom", i, v, 7);
            McRng g;
            mc_rng_seed(&g, ph_seed("lin", tg[i].id));
            for (int k = 0; k < 6; k++) {
                rc = ph_sample_target(&tg[i], &g, &p);
                double w[8] = {k, rc, p.pos.x, p.pos.y, p.pos.z, p.vel.x, p.vel.y, p.vel.z};
                pv("tdraw", i, w, 8);
            }
        }
        for (int i = 0; i < nc; i++) {
            PhaseSpace p;
            int rc = ph_sample_cluster(&cl[i], NULL, &p);
            double v[7] = {rc, p.pos.x, p.pos.y, p.pos.z, p.vel.x, p.vel.y, p.vel.z};
            pv("cnom", i, v, 7);
            McRng g;
            mc_rng_seed(&g, ph_seed("lin", cl[i].name));
            for (int k = 0; k < 3; k++) {
                rc = ph_sample_cluster(&cl[i], &g, &p);
                double w[8] = {k, rc, p.pos.x, p.pos.y, p.pos.z, p.vel.x, p.vel.y, p.vel.z};
                pv("cdraw", i, w, 8);
            }
        }
    } else if (!strcmp(sec, "lininit")) {
        for (int i = 0; i < nt; i++) {
            double ps0[6], sig[6][6];
            ph_lin_target_init(&tg[i], 64, ph_seed("lin", tg[i].id), ps0, sig);
            pv("ps0", i, ps0, 6);
            pv("sig", i, &sig[0][0], 36);
        }
    } else if (!strcmp(sec, "lin")) {
        double tmax = max_tmax(cl, nc);
        for (int i = 0; i < nt; i++) {
            PhLin l;
            ph_lin_target(&tg[i], tmax, 64, ph_seed("lin", tg[i].id), &l);
            print_lin("tlin", i, &l);
            ph_lin_free(&l);
        }
        for (int i = 0; i < nc; i++) {
            PhLin l;
            ph_lin_cluster(&cl[i], 64, ph_seed("lin", cl[i].name), &l);
            print_lin("clin", i, &l);
            ph_lin_free(&l);
        }
    } else if (!strcmp(sec, "chi2")) {
        double tmax = max_tmax(cl, nc);
        PhLin *cn = malloc(sizeof(PhLin) * nc);
        for (int i = 0; i < nc; i++) ph_lin_cluster(&cl[i], 64, ph_seed("lin", cl[i].name), &cn[i]);
        const double rk[2] = {2.0, 0.0};
        for (int r = 0; r < 2; r++) {
            ph_screen_r = rk[r];
            for (int j = 0; j < nt; j++) {
                PhLin tl;
                ph_lin_target(&tg[j], tmax, 64, ph_seed("lin", tg[j].id), &tl);
                for (int i = 0; i < nc; i++) {
                    int kb, kc, kd;
                    double sb, sc, sd;
                    double c0 = ph_screen_chi2(&tl, &cn[i], &cl[i], &kb, &sb);
                    double c1 = ph_screen_chi2_cut(&tl, &cn[i], &cl[i], 25.0, &kc, &sc);
                    double c2 = ph_screen_chi2_cut(&tl, &cn[i], &cl[i], 1e6, &kd, &sd);
                    PhScreen s;
                    ph_screen_finish(&tl, &cn[i], &cl[i], ph_seed(tg[j].id, cl[i].name), c0, kb, sb, &s);
                    double v[17] = {r, i, c0, kb, sb, c1, kc, sc, c2, kd, sd,
                                    s.t, s.d_nom_pc, s.sig_pc, s.p_r50, s.p_rtot, s.vrel};
                    pv("chi2", j, v, 17);
                }
                ph_lin_free(&tl);
            }
        }
        ph_screen_r = 2.0;
    } else if (!strcmp(sec, "pair")) {
        int n_clones = 24;
        double *ps = malloc(sizeof(double) * 12 * n_clones);
        for (int j = 0; j < nt; j++)
            for (int i = 0; i < nc; i++) {
                int steps;
                double dt;
                int m = ph_pair_draw(&tg[j], &cl[i], n_clones, 0.05, ph_seed(tg[j].id, cl[i].name), ps, &steps, &dt);
                double h[4] = {i, m, steps, dt};
                pv("draw", j, h, 4);
                double b[64], t[64], v[64];
                for (int k = 0; k < m; k++) {
                    if (k < 3) pv("clone_ps", j, ps + 12 * k, 12);
                    ph_pair_clone(ps + 12 * k, steps, dt, &b[k], &t[k], &v[k]);
                    double r[4] = {k, b[k], t[k], v[k]};
                    pv("clone", j, r, 4);
                }
                PhResult res;
                ph_pair_reduce(&cl[i], m, b, t, v, &res);
                double o[12] = {i, res.n, res.d05, res.d50, res.d95, res.t05, res.t50, res.t95,
                                res.v05, res.v50, res.v95, res.f_r50};
                pv("reduce", j, o, 12);
                PhResult r2;
                ph_pair(&tg[j], &cl[i], 8, 0.2, ph_seed(tg[j].id, cl[i].name), &r2);
                double o2[13] = {i, r2.n, r2.d05, r2.d50, r2.d95, r2.t05, r2.t50, r2.t95,
                                 r2.v05, r2.v50, r2.v95, r2.f_r50, r2.f_rtot};
                pv("pair", j, o2, 13);
            }
        free(ps);
    } else if (!strcmp(sec, "track")) {
        double tmax = max_tmax(cl, nc);
        PhTrack *ct = malloc(sizeof(PhTrack) * nc);
        for (int i = 0; i < nc; i++) {
            ph_track_cluster(&cl[i], 32, ph_seed("track", cl[i].name), &ct[i]);
            double v[5] = {ct[i].n, ct[i].sig_pos, ct[i].sig_vel, ct[i].pos[ct[i].n - 1].x, ct[i].pos[ct[i].n - 1].y};
            pv("ctrack", i, v, 5);
        }
        for (int j = 0; j < nt; j++) {
            PhTrack tt;
            ph_track_target(&tg[j], tmax, 32, ph_seed("track", tg[j].id), &tt);
            double v[6] = {tt.n, tt.sig_pos, tt.sig_vel, tt.pos[tt.n - 1].x, tt.pos[tt.n - 1].y, tt.pos[tt.n - 1].z};
            pv("ttrack", j, v, 6);
            for (int i = 0; i < nc; i++) printf("prefilter %d %d %d\n", j, i, ph_prefilter(&tt, &ct[i], &cl[i]));
            ph_track_free(&tt);
        }
    } else if (!strcmp(sec, "read")) {
        for (int i = 0; i < nc; i++) {
            double v[17] = {cl[i].ra, cl[i].dec, cl[i].dist_pc, cl[i].e_dist_pc, cl[i].pmra, cl[i].e_pmra,
                            cl[i].pmdec, cl[i].e_pmdec, cl[i].rv, cl[i].e_rv, cl[i].has_rv, cl[i].r50_pc,
                            cl[i].rtot_pc, cl[i].age16, cl[i].age50, cl[i].age84, cl[i].tmax_myr};
            printf("cname %d %s %c\n", i, cl[i].name, cl[i].kind);
            pv("cread", i, v, 17);
        }
        for (int i = 0; i < nt; i++) {
            double v[13] = {tg[i].ra, tg[i].dec, tg[i].plx, tg[i].e_plx, tg[i].dist_pc, tg[i].e_lndist, tg[i].pmra,
                            tg[i].e_pmra, tg[i].pmdec, tg[i].e_pmdec, tg[i].rv, tg[i].e_rv, tg[i].has_rv};
            printf("tname %d %s %c\n", i, tg[i].id, tg[i].kind);
            pv("tread", i, v, 13);
        }
    } else {
        fprintf(stderr, "unknown section %s\n", sec);
        return 1;
    }
    return 0;
}
