/* The reference's core C API, exported by libphygas_hip.so with the same names, signatures, struct layouts and
 * ownership rules, so the reference's command-line programs (src/phygas_screen.c, src/phygas_cli.c) and the CPU
 * loop of its app link against this library in place of phygas.o / potential.o / montecarlo.o.
 *
 * These are the reference's synchronous per-object calls, computed on the host by the library's own FP64
 * physics (phygas_amd/csrc/physics.h, built without FMA contraction and calling the C library's transcendentals,
 * as the reference does), so every result is bit-identical to the reference's on the same platform.  Throughput
 * work belongs on the batched device engine of phygas_hip.h (phx_screen, phx_pair_loaded, phx_pair_ex), which
 * takes the same records and gives the same rows.
 *
 * Units inside: kpc, Myr, kpc/Myr.  Reported: pc, Myr ago, km/s. */
#ifndef PHYGAS_CORE_H
#define PHYGAS_CORE_H

#include <stdint.h>

#include "phygas_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* splitmix64 state (reference: src/montecarlo.h:22-41) */
typedef struct {
    uint64_t state;
} McRng;

static inline void mc_rng_seed(McRng *rng, uint64_t seed) { rng->state = seed; }
static inline uint64_t mc_rng_next_u64(McRng *rng) {
    uint64_t z = (rng->state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double mc_rng_uniform(McRng *rng) { return (double)(mc_rng_next_u64(rng) >> 11) * (1.0 / 9007199254740992.0); }
/* one standard normal per call, Box-Muller (reference: src/montecarlo.h:48, src/montecarlo.c:4-17) */
double mc_gaussian(McRng *rng);

/* potential interface (reference: src/potential.h:13-19) */
typedef struct Potential Potential;
struct Potential {
    void (*acceleration)(const Potential *pot, Vec3 pos, Vec3 *accel_out);
    double (*energy)(const Potential *pot, Vec3 pos);
    const void *params;
};

/* reference: src/phygas.h:64-65.  ph_init may be called any number of times; the library needs no set-up. */
void ph_init(void);
/* the Allen & Santillan 1991 potential of the path (reference: src/potential.c:54-136) */
const Potential *ph_potential(void);

/* reference: src/phygas.h:67-69 -- one draw (rng == NULL: the nominal); 0, or -1 when no draw succeeds */
int ph_sample_target(const PhTarget *t, McRng *rng, PhaseSpace *out);
int ph_sample_cluster(const PhCluster *c, McRng *rng, PhaseSpace *out);

/* reference: src/phygas.h:71-74 -- nominal path at 0.5-Myr points plus its spread from n_spread (<= 64) draws;
 * pos is malloc'ed, free with ph_track_free */
void ph_track_target(const PhTarget *t, double tmax_myr, int n_spread, uint64_t seed, PhTrack *out);
void ph_track_cluster(const PhCluster *c, int n_spread, uint64_t seed, PhTrack *out);
void ph_track_free(PhTrack *t);
/* reference: src/phygas.h:76-78 */
int ph_prefilter(const PhTrack *target, const PhTrack *cluster, const PhCluster *c);

/* reference: src/phygas.h:80-90 -- the Monte Carlo search for one pair and its three parts; ps holds
 * 12 * n_clones doubles (caller's buffer) */
void ph_pair(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, PhResult *res);
int ph_pair_draw(const PhTarget *t, const PhCluster *c, int n_clones, double dt_myr, uint64_t seed, double *ps,
                 int *steps, double *dt);
void ph_pair_clone(const double *ps, int steps, double dt, double *best, double *tbest, double *vbest);
void ph_pair_reduce(const PhCluster *c, int m, const double *best, const double *tbest, const double *vbest,
                    PhResult *res);

/* reference: src/phygas.h:116-133 -- linear tracks (pos, vel, S malloc'ed; free with ph_lin_free) and the
 * screening of one pair */
void ph_lin_target(const PhTarget *t, double tmax_myr, int n_samp, uint64_t seed, PhLin *out);
void ph_lin_target_init(const PhTarget *t, int n_samp, uint64_t seed, double ps0[6], double sig[6][6]);
void ph_lin_cluster(const PhCluster *c, int n_samp, uint64_t seed, PhLin *out);
void ph_lin_free(PhLin *l);
extern double ph_screen_r; /* kernel radius in units of r50 (default 2); 0 = the total radius */
void ph_screen(const PhLin *t, const PhLin *c, const PhCluster *cl, uint64_t seed, PhScreen *res);
double ph_screen_chi2(const PhLin *t, const PhLin *c, const PhCluster *cl, int *kb, double *sb);
double ph_screen_chi2_cut(const PhLin *t, const PhLin *c, const PhCluster *cl, double chi2_max, int *kb, double *sb);
void ph_screen_finish(const PhLin *t, const PhLin *c, const PhCluster *cl, uint64_t seed, double chi2, int kb,
                      double sb, PhScreen *res);

/* reference: src/phygas.h:135 */
uint64_t ph_seed(const char *target_id, const char *cluster_name);

/* reference: src/phygas.h:137-139 -- records malloc'ed into *out (free()); count, -1 when the file cannot be
 * opened; lines that do not parse are skipped */
int ph_read_clusters(const char *path, PhCluster **out);
int ph_read_targets(const char *path, PhTarget **out);

#ifdef __cplusplus
}
#endif
#endif
