/* ORACLE -- test infrastructure only.
 *
 * CPU FP64 restatement of the Phygas trace-back hot path (reference: src/phygas.c, src/potential.c,
 * src/montecarlo.[ch], src/coords.c).  It exports the reference's core API under the reference's own
 * names and struct layouts (reference: src/phygas.h:64-139), so the golden harness compiles
 * unchanged against either, and a set of orc_* batch drivers for the tests and the CPU baseline.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and
 * only as the checker: the product (phygas_amd/) never links or calls it. */
#ifndef PHYGAS_ORACLE_H
#define PHYGAS_ORACLE_H

#include <stdint.h>
#include "phygas_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- reference core API (same names and signatures as src/phygas.h:64-139) --- */
typedef struct { uint64_t state; } McRng;
void mc_rng_seed(McRng *rng, uint64_t seed);
uint64_t mc_rng_next_u64(McRng *rng);
double mc_rng_uniform(McRng *rng);
double mc_gaussian(McRng *rng);

/* Potential interface (reference: src/potential.h:15-25) */
typedef struct Potential Potential;
struct Potential {
    void (*acceleration)(const Potential *pot, Vec3 pos, Vec3 *accel_out);
    double (*energy)(const Potential *pot, Vec3 pos);
    const void *params;
};

void ph_init(void);
const Potential *ph_potential(void);
int ph_sample_target(const PhTarget *t, McRng *rng, PhaseSpace *out);
int ph_sample_cluster(const PhCluster *c, McRng *rng, PhaseSpace *out);
void ph_track_target(const PhTarget *t, double tmax_myr, int n_spread, uint64_t seed, PhTrack *out);
void ph_track_cluster(const PhCluster *c, int n_spread, uint64_t seed, PhTrack *out);
void ph_track_free(PhTrack *t);
int ph_prefilter(const PhTrack *target, const PhTrack *cluster, const PhCluster *c);
void ph_pair(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, PhResult *res);
int ph_pair_draw(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed,
                 double *ps, int *steps, double *dt);
void ph_pair_clone(const double *ps, int steps, double dt, double *best, double *tbest, double *vbest);
void ph_pair_reduce(const PhCluster *c, int m, const double *best, const double *tbest, const double *vbest,
                    PhResult *res);
void ph_lin_target(const PhTarget *t, double tmax_myr, int n_samp, uint64_t seed, PhLin *out);
void ph_lin_target_init(const PhTarget *t, int n_samp, uint64_t seed, double ps0[6], double sig[6][6]);
void ph_lin_cluster(const PhCluster *c, int n_samp, uint64_t seed, PhLin *out);
void ph_lin_free(PhLin *l);
extern double ph_screen_r;
void ph_screen(const PhLin *t, const PhLin *c, const PhCluster *cl, uint64_t seed, PhScreen *res);
double ph_screen_chi2(const PhLin *t, const PhLin *c, const PhCluster *cl, int *kb, double *sb);
double ph_screen_chi2_cut(const PhLin *t, const PhLin *c, const PhCluster *cl, double chi2_max, int *kb,
                          double *sb);
void ph_screen_finish(const PhLin *t, const PhLin *c, const PhCluster *cl, uint64_t seed, double chi2, int kb,
                      double sb, PhScreen *res);
uint64_t ph_seed(const char *target_id, const char *cluster_name);
int ph_read_clusters(const char *path, PhCluster **out);
int ph_read_targets(const char *path, PhTarget **out);

/* --- oracle-only helpers (no reference counterpart) --- */
/* Acceleration of the Allen & Santillan potential at (x,y,z), and its potential energy. */
void orc_accel(double x, double y, double z, double out[3]);
double orc_energy(double x, double y, double z);
/* One kick-drift-kick step of b = {pos, vel, acc}. */
void orc_body_step(double b[9], double dt);
/* Observed astrometry -> Galactocentric phase space and back (reference: src/coords.c:77-161). */
void orc_astrometry_to_gc(double ra, double dec, double plx, double pmra, double pmdec, double rv, double out[6]);
void orc_gc_to_astrometry(const double ps[6], double out[6]);
/* symmetric 3x3 m^T M^-1 m (-1 if singular) and the closest point of a segment to the origin */
double orc_quad_inv(const double M[6], const double m[3]);
double orc_seg_min(const double r0[3], const double r1[3], double *s);
/* Potential parameters in device order (reference: src/main_phygas_boinc.cpp:79-80). */
void orc_params(double p[9]);

/* One screened pair as a flat record. */
typedef struct {
    int star, cluster, kb, pad;
    double chi2, sb, t, d_nom_pc, sig_pc, p_r50, p_rtot, vrel;
} OrcRow;

/* Linear tracks of every cluster / target into flat arrays: pos,vel [n][3], S [n][6] per object at
 * offset off[i] (points); n[i] points each.  Returns total points or -1. */
long orc_lin_clusters(const PhCluster *cl, int nc, int n_samp, int nthreads, int *n, long *off, double *pos,
                      double *vel, double *S, long cap);
long orc_lin_targets(const PhTarget *tg, int nt, double tmax, int n_samp, int nthreads, int *n, long *off,
                     double *pos, double *vel, double *S, long cap);
/* Star draws: nominal start state and the 6x6 covariance of the draws (ph_lin_target_init). */
void orc_lin_init_targets(const PhTarget *tg, int nt, int n_samp, double *ps0, double *sig);

/* Full stage-1 screening of targets [t0, t1) against every cluster, as phygas_screen does it (cut at
 * chi2_max, then the 512-draw finish).  use_cut = 0 runs the plain full scan.  Rows in (target,
 * cluster) order into rows[cap]; returns the row count (or -(needed) when cap is too small). */
long orc_screen_run(const PhCluster *cl, int nc, const PhTarget *tg, int nt, int t0, int t1, double chi2_max,
                    int n_samp, double r_kernel, int use_cut, int nthreads, OrcRow *rows, long cap);
/* Per-pair chi2 of every (target, cluster) with no cut (for kernel parity of the full scan). */
void orc_chi2_matrix(const PhCluster *cl, int nc, const PhTarget *tg, int nt, int n_samp, double r_kernel,
                     int nthreads, double *chi2, int *kb, double *sb);
/* Monte Carlo pairs: results of ph_pair for each (target ti[i], cluster ci[i]). */
void orc_pair_run(const PhTarget *tg, const PhCluster *cl, int npairs, const int *ti, const int *ci, int n_clones,
                  double dt_myr, int nthreads, PhResult *res);
/* Clone results (best kpc, tbest Myr, vbest kpc/Myr) of pre-drawn clones ps[n][12]. */
void orc_pair_clones(const double *ps, const int *steps, const double *dts, int n, int nthreads, double *best,
                     double *tbest, double *vbest);

/* Pair-mode prefilter of every (target, cluster) as phygas_cli runs it: mask[j * nc + i]. */
void orc_prefilter_mask(const PhTarget *tg, int nt, const PhCluster *cl, int nc, unsigned char *mask);

#ifdef __cplusplus
}
#endif
#endif
