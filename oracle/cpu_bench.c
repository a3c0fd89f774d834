/* CPU baseline timer -- test/bench infrastructure.
 *
 * Built by oracle/Makefile against the compiled reference (oracle/_ref/cpu_bench_ref) and against the oracle
 * (oracle/cpu_bench_orc).  Every call it times is reentrant once ph_init has run; work is spread over threads
 * by star (screening) or by pair (Monte Carlo).  Configurations run interleaved, round after round, so slow
 * drifts of the machine hit every configuration alike; each configuration's runs are summarised by their
 * median, minimum and maximum (the reference's bench method, scripts/bench/summarize_bench.py).
 *
 *   cpu_bench screen clusters.txt stars.txt reps T:N [T:N ...]
 *       the stage-1 loop of phygas_screen (reference: src/phygas_screen.c:48-67: per star ph_lin_target, then
 *       ph_screen_chi2_cut and ph_screen_finish for every cluster) over the first N stars on T threads; the
 *       cluster tracks are prepared once first (timed apart: they stay resident on the GPU)
 *   cpu_bench pair clusters.txt targets.txt reps n_clones dt T:N [T:N ...]
 *       ph_pair of target 0 against clusters 0 .. N-1 on T threads (reference: src/phygas.c:240-249)
 * One line per configuration: key=value pairs. */
#define _POSIX_C_SOURCE 200809L
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef USE_REF
#include "phygas.h"
#else
#include "phygas_oracle.h"
#endif

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static PhCluster *g_cl;
static PhTarget *g_tg;
static PhLin *g_lin;
static int g_nc, g_nwork, g_next, g_nsamp = 64, g_clones = 200;
static double g_tmax, g_chi2_max = 25.0, g_dt = 0.05;
static long g_kept, g_rows_p001;
static double g_steps;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

static int next_item(void) {
    pthread_mutex_lock(&g_mu);
    int i = g_next++;
    pthread_mutex_unlock(&g_mu);
    return i;
}

static void *cluster_worker(void *arg) {
    (void)arg;
    for (int i; (i = next_item()) < g_nc;) ph_lin_cluster(&g_cl[i], g_nsamp, ph_seed("lin", g_cl[i].name), &g_lin[i]);
    return NULL;
}

static void *star_worker(void *arg) {
    (void)arg;
    long kept = 0, rows = 0;
    for (int j; (j = next_item()) < g_nwork;) {
        PhLin tl;
        ph_lin_target(&g_tg[j], g_tmax, g_nsamp, ph_seed("lin", g_tg[j].id), &tl);
        for (int i = 0; i < g_nc; i++) {
            int kb;
            double sb, chi2 = ph_screen_chi2_cut(&tl, &g_lin[i], &g_cl[i], g_chi2_max, &kb, &sb);
            if (!(chi2 < g_chi2_max)) continue;
            PhScreen s;
            ph_screen_finish(&tl, &g_lin[i], &g_cl[i], ph_seed(g_tg[j].id, g_cl[i].name), chi2, kb, sb, &s);
            kept++;
            if (s.p_r50 >= 0.001) rows++;
        }
        ph_lin_free(&tl);
    }
    pthread_mutex_lock(&g_mu);
    g_kept += kept;
    g_rows_p001 += rows;
    pthread_mutex_unlock(&g_mu);
    return NULL;
}

static void *pair_worker(void *arg) {
    (void)arg;
    double steps = 0;
    for (int i; (i = next_item()) < g_nwork;) {
        PhResult r;
        ph_pair(&g_tg[0], &g_cl[i], g_clones, g_dt, ph_seed(g_tg[0].id, g_cl[i].name), &r);
        steps += (double)r.n * ceil(g_cl[i].tmax_myr / g_dt);
    }
    pthread_mutex_lock(&g_mu);
    g_steps += steps;
    pthread_mutex_unlock(&g_mu);
    return NULL;
}

static double run(void *(*fn)(void *), int nth) {
    pthread_t th[256];
    g_next = 0;
    double t0 = now_s();
    for (int t = 0; t < nth; t++) pthread_create(&th[t], NULL, fn, NULL);
    for (int t = 0; t < nth; t++) pthread_join(th[t], NULL);
    return now_s() - t0;
}

static int cmp_d(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "usage: %s screen clusters stars reps T:N [T:N ...]\n"
                        "       %s pair clusters targets reps n_clones dt T:N [T:N ...]\n", argv[0], argv[0]);
        return 1;
    }
    const int pair = strcmp(argv[1], "pair") == 0;
    ph_init();
    g_nc = ph_read_clusters(argv[2], &g_cl);
    int nt = ph_read_targets(argv[3], &g_tg);
    if (g_nc <= 0 || nt <= 0) { fprintf(stderr, "can't read inputs\n"); return 1; }
    int reps = atoi(argv[4]);
    if (reps < 1) reps = 1;
    if (reps > 64) reps = 64;
    int a0 = 5;
    if (pair) {
        if (argc < 8) { fprintf(stderr, "pair mode needs n_clones, dt and T:N\n"); return 1; }
        g_clones = atoi(argv[5]);
        g_dt = atof(argv[6]);
        a0 = 7;
    }
    int ncfg = argc - a0, th[32], nw[32], maxth = 1;
    if (ncfg > 32) ncfg = 32;
    for (int c = 0; c < ncfg; c++) {
        if (sscanf(argv[a0 + c], "%d:%d", &th[c], &nw[c]) != 2) { fprintf(stderr, "bad T:N\n"); return 1; }
        if (th[c] < 1) th[c] = 1;
        if (th[c] > 256) th[c] = 256;
        if (nw[c] > (pair ? g_nc : nt)) nw[c] = pair ? g_nc : nt;
        if (th[c] > maxth) maxth = th[c];
    }
    double prep = 0;
    if (!pair) {
        for (int i = 0; i < g_nc; i++) if (g_cl[i].tmax_myr > g_tmax) g_tmax = g_cl[i].tmax_myr;
        g_lin = malloc(sizeof(PhLin) * g_nc);
        prep = run(cluster_worker, maxth);
    }
    static double t[32][64];
    long kept[32] = {0}, rows[32] = {0};
    double steps[32] = {0};
    for (int r = 0; r < reps; r++)
        for (int c = 0; c < ncfg; c++) {
            g_nwork = nw[c];
            g_kept = g_rows_p001 = 0;
            g_steps = 0;
            t[c][r] = run(pair ? pair_worker : star_worker, th[c]);
            kept[c] = g_kept;
            rows[c] = g_rows_p001;
            steps[c] = g_steps;
        }
    for (int c = 0; c < ncfg; c++) {
        qsort(t[c], reps, sizeof(double), cmp_d);
        const double med = reps % 2 ? t[c][reps / 2] : 0.5 * (t[c][reps / 2 - 1] + t[c][reps / 2]);
        if (pair)
            printf("mode=pair threads=%d pairs=%d clones=%d clone_steps=%.0f runs=%d median_s=%.6f min_s=%.6f "
                   "max_s=%.6f clone_steps_per_s=%.6e\n",
                   th[c], nw[c], g_clones, steps[c], reps, med, t[c][0], t[c][reps - 1], steps[c] / med);
        else
            printf("mode=screen threads=%d stars=%d clusters=%d pairs=%ld kept=%ld rows_p001=%ld cluster_prep_s=%.6f "
                   "runs=%d median_s=%.6f min_s=%.6f max_s=%.6f pairs_per_s=%.6e\n",
                   th[c], nw[c], g_nc, (long)nw[c] * g_nc, kept[c], rows[c], prep, reps, med, t[c][0],
                   t[c][reps - 1], (double)nw[c] * g_nc / med);
    }
    return 0;
}
