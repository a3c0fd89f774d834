/* CPU baseline timer -- test/bench infrastructure.
 *
 * Runs the stage-1 screening loop of phygas_screen (reference: src/phygas_screen.c:48-67: per star
 * ph_lin_target, then ph_screen_chi2_cut and ph_screen_finish for every cluster) over the first
 * n_stars targets, with the work spread over nthreads threads by star (every call is reentrant once
 * ph_init has run; the kernel radius is the process-wide ph_screen_r, left at its default 2).
 * The cluster tracks are prepared first and timed separately, as they are resident on the GPU.
 *
 * Built by oracle/Makefile against the compiled reference (oracle/_ref/cpu_bench_ref) and against
 * the oracle (oracle/cpu_bench_orc).  Prints one line: key=value pairs.
 *   cpu_bench clusters.txt stars.txt nthreads n_stars [chi2_max] */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
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
static int g_nc, g_ns, g_next, g_nsamp = 64;
static double g_tmax, g_chi2_max = 25.0;
static long g_kept, g_rows_p001;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

static void *cluster_worker(void *arg) {
    (void)arg;
    for (;;) {
        pthread_mutex_lock(&g_mu);
        int i = g_next++;
        pthread_mutex_unlock(&g_mu);
        if (i >= g_nc) break;
        ph_lin_cluster(&g_cl[i], g_nsamp, ph_seed("lin", g_cl[i].name), &g_lin[i]);
    }
    return NULL;
}

static void *star_worker(void *arg) {
    (void)arg;
    long kept = 0, rows = 0;
    for (;;) {
        pthread_mutex_lock(&g_mu);
        int j = g_next++;
        pthread_mutex_unlock(&g_mu);
        if (j >= g_ns) break;
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

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s clusters.txt stars.txt nthreads n_stars [chi2_max]\n", argv[0]);
        return 1;
    }
    int nth = atoi(argv[3]);
    if (nth < 1) nth = 1;
    if (nth > 256) nth = 256;
    if (argc > 5) g_chi2_max = atof(argv[5]);
    ph_init();
    int nt;
    g_nc = ph_read_clusters(argv[1], &g_cl);
    nt = ph_read_targets(argv[2], &g_tg);
    if (g_nc <= 0 || nt <= 0) { fprintf(stderr, "can't read inputs\n"); return 1; }
    g_ns = atoi(argv[4]);
    if (g_ns > nt) g_ns = nt;
    for (int i = 0; i < g_nc; i++) if (g_cl[i].tmax_myr > g_tmax) g_tmax = g_cl[i].tmax_myr;
    g_lin = malloc(sizeof(PhLin) * g_nc);
    pthread_t th[256];
    double t0 = now_s();
    g_next = 0;
    for (int t = 0; t < nth; t++) pthread_create(&th[t], NULL, cluster_worker, NULL);
    for (int t = 0; t < nth; t++) pthread_join(th[t], NULL);
    double t1 = now_s();
    g_next = 0;
    for (int t = 0; t < nth; t++) pthread_create(&th[t], NULL, star_worker, NULL);
    for (int t = 0; t < nth; t++) pthread_join(th[t], NULL);
    double t2 = now_s();
    long pairs = (long)g_ns * g_nc;
    printf("threads=%d stars=%d clusters=%d pairs=%ld kept=%ld rows_p001=%ld cluster_prep_s=%.6f screen_s=%.6f "
           "pairs_per_s=%.6e\n",
           nth, g_ns, g_nc, pairs, g_kept, g_rows_p001, t1 - t0, t2 - t1, pairs / (t2 - t1));
    return 0;
}
