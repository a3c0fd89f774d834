/* Record types of the Phygas trace-back path, layout-compatible with the reference's core API
 * (reference: src/phygas.h:19-57 PhCluster/PhTarget/PhTrack/PhResult, :100-113 PhLin/PhScreen,
 * src/vec3.h:6 Vec3, src/coords.h:21-24 PhaseSpace).  Plain C99; no device types.
 *
 * Units inside: kpc, Myr, kpc/Myr.  Reported: pc, Myr ago (positive), km/s. */
#ifndef PHYGAS_TYPES_H
#define PHYGAS_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y, z; } Vec3;

typedef struct {
    Vec3 pos; /* kpc, Galactocentric */
    Vec3 vel; /* kpc/Myr */
} PhaseSpace;

/* One young cluster or moving group: one line of the 19-field cluster file. */
typedef struct {
    char name[40];
    char kind;                 /* o = open cluster, m = moving group */
    double ra, dec;            /* deg */
    double dist_pc, e_dist_pc;
    double pmra, e_pmra, pmdec, e_pmdec; /* mas/yr, pmra includes cos dec */
    double rv, e_rv;           /* km/s */
    int has_rv;                /* 0: drawn around the rotation expectation +- 10 km/s */
    double r50_pc, rtot_pc;
    double age16, age50, age84; /* Myr */
    double tmax_myr;           /* trace-back length */
} PhCluster;

/* One target (star or pulsar): one line of the 14-field target file. */
typedef struct {
    char id[40];
    char kind;                 /* P = pulsar / neutron star, S = star */
    double ra, dec;
    double plx, e_plx;         /* mas; used when plx > 0 */
    double dist_pc, e_lndist;  /* used when plx == 0: log-normal distance */
    double pmra, e_pmra, pmdec, e_pmdec;
    double rv, e_rv;
    int has_rv;
} PhTarget;

/* Nominal path plus spread (pair-mode prefilter). */
typedef struct {
    int n;
    Vec3 *pos;
    double sig_pos, sig_vel;
} PhTrack;

/* Monte Carlo result of one pair. */
typedef struct {
    int n;
    double d05, d50, d95;  /* pc */
    double t05, t50, t95;  /* Myr ago */
    double v05, v50, v95;  /* km/s */
    double f_r50, f_rtot;
} PhResult;

/* Linearised track: nominal pos/vel and position covariance (xx xy xz yy yz zz) per 0.5-Myr point. */
typedef struct {
    int n;
    Vec3 *pos, *vel;
    double (*S)[6];
} PhLin;

/* Screening result of one pair. */
typedef struct {
    double chi2, t, d_nom_pc, sig_pc, p_r50, p_rtot, vrel;
} PhScreen;

#define PH_PRE_STEP_MYR 0.5
#define PH_CLUSTER_RV_PRIOR 10.0
#define PH_PULSAR_RV_PRIOR 265.0
#define PH_STAR_RV_PRIOR 30.0
#define PH_LIN_EPS_POS 1e-5
#define PH_LIN_EPS_VEL 1e-5
#define PH_LIN_SUB 10

#ifdef __cplusplus
}
#endif
#endif
